# Run TDGL 2.0 with the HIP ladder cycle; measurements and checkpoints stay on the host (as run_gpu.jl).
#   julia --project=julia/TDGLHIP julia/TDGLHIP/scripts/run_hip.jl RUN.toml [--dir OUT] [--resume] [--polarity +|-] [--require-gpu]
using TDGL, TDGLHIP, Printf
isempty(ARGS) && (println("usage: run_hip.jl RUN.toml [--dir OUT] [--resume] [--polarity +|-] [--require-gpu]"); exit(0))
opts = Dict{String,String}(); flags = Set{String}(); i = 2
while i <= length(ARGS)
    a = ARGS[i]
    if a in ("--dir", "--polarity"); opts[a] = ARGS[i+1]; i += 2
    elseif a in ("--resume", "--require-gpu"); push!(flags, a); i += 1
    else error("unknown option: $a") end
end
p = get(opts, "--polarity", "both")
pol = p == "both" ? :both : p in ("+", "pos", "J+") ? :pos : :neg
TDGLHIP.enable!(; fallback="--require-gpu" in flags ? :error : :cpu)
t0 = time()
r = try
    run_config(ARGS[1]; out_dir=get(opts, "--dir", nothing), resume="--resume" in flags, polarity=pol)
finally
    TDGLHIP.disable!()
end
st = execution_stats()
@printf("wall %.1f min; backend=%s, backend_steps=%d, cpu_fallback_steps=%d\n", (time() - t0) / 60, st.backend,
        st.backend_steps, st.cpu_fallback_steps)
