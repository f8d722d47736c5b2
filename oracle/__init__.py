"""CPU oracle of the release-2.0 step.  TEST INFRASTRUCTURE: only tests/, the smoke check and the
benchmark's cpu_baseline leg use it, as the checker; the product (juliatdgl_amd) never does."""
