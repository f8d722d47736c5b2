# ============================================================================
#  TDGLHIP — the ladder cycle of TDGL 2.0 on an MI355X through the C ABI of libtdglhip
#  (include/tdglhip.h).  The same exports as TDGLGPU (gpu/TDGLGPU/src/TDGLGPU.jl:29), so that
#  run_gpu.jl-style drivers switch by `TDGLHIP.enable!()` in place of `TDGLGPU.enable!(backend)`.
#
#      gs = GPUStepper(st)            # st: a TDGL.Stepper the HIP stepper runs (supported(st))
#      g  = device_state(gs, s)       # s: a TDGL.State
#      gpu_steps!(g, gs, n; H, HJ)    # n steps on the device
#      host_state!(s, g)              # back into s
#
#  The library path: ENV["TDGLHIP_LIB"], else juliatdgl_amd/_lib/libtdglhip.so of this repository.
# ============================================================================
module TDGLHIP

using Libdl
using TDGL
using TDGL: State, Stepper, SepSolver, PCGSolver

export GPUStepper, GState, device_state, device_state!, host_state!, gpu_step!, gpu_steps!, supported, enable!,
       meter_begin!, meter_cycle, noise_counter, noise_counter!, noise_draws!, reference_draws, noise_position,
       noise_position!,
       disable!, execution_stats

const MAXB = 4
const MAXS = 64
const _LIBPATH = Ref{String}("")
libpath() = isempty(_LIBPATH[]) ? get(ENV, "TDGLHIP_LIB",
    normpath(joinpath(@__DIR__, "..", "..", "..", "juliatdgl_amd", "_lib", "libtdglhip.so"))) : _LIBPATH[]

# ---- the C structs (include/tdglhip.h), field for field -------------------------------------
struct BandCoefC
    kx::Float64; ky::Float64; kz::Float64; c::Float64; eps::Float64; beta::Float64; gamma2::Float64
    cpl::NTuple{MAXB,Float64}
end
struct DescC
    abi_version::Int32; nbands::Int32
    Nx::Int32; Ny::Int32; Nz::Int32
    x1::Int32; x2::Int32; y1::Int32; y2::Int32; z1::Int32; z2::Int32
    dx::Float64; dy::Float64; dz::Float64; dt::Float64
    bc::NTuple{MAXB,BandCoefC}; bcs::NTuple{MAXB,BandCoefC}
    nsub::NTuple{MAXB,Int32}; stages::NTuple{MAXB,Int32}; rkc::NTuple{MAXB,Int32}
    divisor::NTuple{MAXB,Float64}
    rkc_mu::NTuple{MAXB,NTuple{MAXS,Float64}}; rkc_nu::NTuple{MAXB,NTuple{MAXS,Float64}}
    rkc_c::NTuple{MAXB,NTuple{MAXS,Float64}}
    wa::NTuple{3,NTuple{MAXB,Float64}}
    d1::Float64; sxy::Float64; sxz::Float64; syx::Float64; syz::Float64; szx::Float64; szy::Float64; a::Float64
    field_mode::Int32; selffield_bc::Int32
    extrap::Int32; rotframe::Int32; hard::Int32
    pcg_tol::Float64; pcg_maxit::Int32; pcg_fast::Int32; pcg_f32::Int32; pcg_extrap::Int32
    f::Ptr{Float64}; taux::Ptr{Float64}; tauy::Ptr{Float64}; tauz::Ptr{Float64}
    Dx::Ptr{Float64}; Dy::Ptr{Float64}; Dz::Ptr{Float64}
    bar_hy_z::Ptr{Float64}; bar_hy_y::Ptr{Float64}; bar_hz_z::Ptr{Float64}; bar_hz_y::Ptr{Float64}
    bar_hz_x::Ptr{Float64}; bar_hy_x::Ptr{Float64}
    # normal fluids (exchange.jl)
    qp_mode::Int32; qp_nmodes::Int32
    qp_j0::Int32; qp_nj::Int32; qp_k0::Int32; qp_nk::Int32
    qp_sigma::Float64
    qp_s::NTuple{MAXB,Float64}; qp_turn::NTuple{MAXB,Float64}; qp_lambda::NTuple{MAXB,Float64}
    qp_P::NTuple{MAXB,NTuple{MAXB,Float64}}; qp_Pinv::NTuple{MAXB,NTuple{MAXB,Float64}}
    qp_gamma::NTuple{MAXB,NTuple{MAXB,Float64}}
    qp_Vx::Ptr{Float64}; qp_Wx::Ptr{Float64}; qp_ex::Ptr{Float64}
    qp_Vy::Ptr{Float64}; qp_Wy::Ptr{Float64}; qp_ey::Ptr{Float64}
    qp_Vz::Ptr{Float64}; qp_Wz::Ptr{Float64}; qp_ez::Ptr{Float64}
    # open boundary (openbc.jl)
    open_every::Int32; open_R::Int32; open_fac::Float64; open_kappa2::Float64
    open_lgf::Ptr{Float64}; open_Jbar::Ptr{Float64}
    # thermal noise (noise.jl)
    noise::Int32; noise_field::Int32; noise_host::Int32; noise_sk::NTuple{MAXB,Int32}
    noise_T::Float64; noise_seed::UInt64
    noise_sk_w::NTuple{MAXB,NTuple{2,Float64}}; noise_Gamma::NTuple{MAXB,Float64}; noise_sigma_n::Float64
    # the PCG's precision (an extension: FP64 refinement around Float32 inner solves)
    pcg_refine::Int32; pcg_rounds::Int32; pcg_inner_tol::Float64
end
struct StateC
    G::NTuple{MAXB,Ptr{Float64}}; F::NTuple{MAXB,Ptr{Float64}}
    thx::Ptr{Float64}; thy::Ptr{Float64}; thz::Ptr{Float64}
    phx::Ptr{Float64}; phy::Ptr{Float64}; phz::Ptr{Float64}
    Ux::Ptr{Float64}; Uy::Ptr{Float64}; Uz::Ptr{Float64}
    Vx::Ptr{Float64}; Vy::Ptr{Float64}; Vz::Ptr{Float64}
    Jx::Ptr{Float64}; Jy::Ptr{Float64}; Jz::Ptr{Float64}
    hx::Ptr{Float64}; hy::Ptr{Float64}; hz::Ptr{Float64}
end
struct PcgC
    x::NTuple{3,Ptr{Float64}}; xprev::NTuple{3,Ptr{Float64}}
    warm::Int32; have2::Int32; iters::Int32
end

_pad(t, n, z) = ntuple(i -> i <= length(t) ? t[i] : z, n)
_bc(b, N, µ) = BandCoefC(b.kx, b.ky, b.kz, b.c, b.ε, b.β, b.γ2, _pad(ntuple(ν -> ν == µ ? 0.0 : b.cpl[ν], N), MAXB, 0.0))
const _ZBC = BandCoefC(0, 0, 0, 0, 0, 0, 0, ntuple(_ -> 0.0, MAXB))

"""
    supported(st) -> (ok::Bool, why::String)

Whether the HIP stepper runs the step of this Stepper (TDGLGPU.jl:39-44), and if not the first
option it does not have.
"""
function supported(st::Stepper{N}) where N
    N <= MAXB || return (false, "more than $MAXB bands")
    st.heat === nothing || return (false, "the heat model")
    st.drag === nothing || return (false, "interband drag")
    st.mt === nothing || return (false, "materials per region")
    if st.qpx !== nothing
        st.qpx.ps.cond === nothing || return (false, "normal fluids in a sculpted sample")
        st.qpx isa TDGL.QPExchange && st.qpx.var !== nothing && return (false, "normal fluids with materials")
    end
    st.masks.vac === nothing || return (false, "a sculpted sample")
    any(>(MAXS), st.rkl) && return (false, "more than $MAXS stages")
    (true, "")
end

# the description of st, and the arrays it points to (kept alive by the caller)
function _desc(st::Stepper{N}) where N
    gr = st.grid; mk = st.masks; fc = st.fc
    keep = Any[]
    hold(a) = (b = Array{Float64}(a); push!(keep, b); pointer(b))
    mode = st.imp isa PCGSolver ? 2 : st.imp isa SepSolver ? 1 : st.D !== nothing ? 3 : 0
    D = mode == 2 ? st.imp.D : mode == 3 ? st.D : nothing
    pcg = mode == 2 ? st.imp : nothing
    fast = pcg !== nothing && pcg.fast !== nothing
    rk = ntuple(µ -> st.rkc !== nothing && st.rkc[µ] !== nothing ? st.rkc[µ] : nothing, N)
    div = ntuple(N) do µ
        st.rkl[µ] > 0 ? (rk[µ] !== nothing ? rk[µ].q : Float64(st.rkl[µ] * (st.rkl[µ] + 1) ÷ 2)) : Float64(st.nsub[µ])
    end
    tab(f) = _pad(ntuple(µ -> rk[µ] === nothing ? ntuple(_ -> 0.0, MAXS) : _pad(Tuple(f(rk[µ])), MAXS, 0.0), N),
                  MAXB, ntuple(_ -> 0.0, MAXS))
    bar = st.bar
    bp(k) = bar === nothing ? Ptr{Float64}(0) : hold(getproperty(bar, k))
    d = DescC(5, N, gr.Nx, gr.Ny, gr.Nz, gr.x1, gr.x2, gr.y1, gr.y2, gr.z1, gr.z2, gr.dx, gr.dy, gr.dz, st.dt,
              _pad(ntuple(µ -> _bc(st.bc[µ], N, µ), N), MAXB, _ZBC), _pad(ntuple(µ -> _bc(st.bcs[µ], N, µ), N), MAXB, _ZBC),
              _pad(Int32.(st.nsub), MAXB, Int32(0)), _pad(Int32.(st.rkl), MAXB, Int32(0)),
              _pad(ntuple(µ -> Int32(rk[µ] !== nothing), N), MAXB, Int32(0)), _pad(div, MAXB, 0.0),
              tab(k -> k.μ), tab(k -> k.ν), tab(k -> k.c),
              ntuple(i -> _pad(st.wa[i], MAXB, 0.0), 3),
              fc.d1, fc.sxy, fc.sxz, fc.syx, fc.syz, fc.szx, fc.szy, fc.a,
              mode, st.open !== nothing ? 2 : bar === nothing ? 0 : 1, st.extrap, st.rotframe, st.hard,
              pcg === nothing ? 0.0 : pcg.tol, pcg === nothing ? 0 : pcg.maxit,
              Int32(fast), Int32(fast && pcg.fast.f32), Int32(fast && pcg.fast.extrap),
              hold(mk.f), hold(mk.τx), hold(mk.τy), hold(mk.τz),
              D === nothing ? Ptr{Float64}(0) : hold(D[1]), D === nothing ? Ptr{Float64}(0) : hold(D[2]),
              D === nothing ? Ptr{Float64}(0) : hold(D[3]),
              bp(:hy_z), bp(:hy_y), bp(:hz_z), bp(:hz_y), bp(:hz_x), bp(:hy_x),
              _qp_fields(st, hold)..., _open_fields(st, hold)..., _noise_fields(st, rk)...,
              Int32(PCG_REFINE[] && fast && pcg.fast.f32), Int32(PCG_ROUNDS[]), PCG_INNER_TOL[])
    d, keep
end

# The PCG's precision on the device (an extension of the reference's solver): with PCG_REFINE[] the
# FP64 operator, residual and solution refine around Float32 inner solves, each to PCG_INNER_TOL[]
# of its starting residual, PCG_ROUNDS[] rounds per captured solve.  Read when a handle is created.
const PCG_REFINE = Ref(false)
const PCG_ROUNDS = Ref(3)
const PCG_INNER_TOL = Ref(1e-4)

# the Noise of st (noise.jl:32-41) as the description's noise fields: draws on the device unless
# HOST_DRAWS[] (then the hook hands over the reference's own stream, reference_draws)
const HOST_DRAWS = Ref(false)
function _noise_fields(st::Stepper{N}, rk) where N
    nz = st.noise
    nz === nothing && return (Int32(0), Int32(0), Int32(0), ntuple(_ -> Int32(0), MAXB), 0.0, UInt64(0),
                              ntuple(_ -> (0.0, 0.0), MAXB), _Z4, 0.0)
    sk = _pad(ntuple(µ -> Int32(TDGL._skrock(st, µ)), N), MAXB, Int32(0))
    w = _pad(ntuple(µ -> rk[µ] === nothing ? (0.0, 0.0) : rk[µ].sk, N), MAXB, (0.0, 0.0))
    Γ = _pad(ntuple(µ -> st.model.bands[µ].Γ, N), MAXB, 0.0)
    (Int32(1), Int32(nz.field), Int32(HOST_DRAWS[]), sk, nz.T, nz.seed, w, Γ, st.model.σn)
end

const _Z4 = ntuple(_ -> 0.0, MAXB)
const _Z44 = ntuple(_ -> _Z4, MAXB)
_null() = Ptr{Float64}(0)

# the QPExchange / QPExchangeN of st as the description's qp fields (exchange.jl:130-146, 425-444)
function _qp_fields(st::Stepper{N}, hold) where N
    q = st.qpx
    q === nothing && return (Int32(0), Int32(0), Int32(0), Int32(0), Int32(0), Int32(0), 0.0, _Z4, _Z4, _Z4, _Z44, _Z44,
                             _Z44, ntuple(_ -> _null(), 9)...)
    _, nj, nk = q isa TDGL.QPExchange ? size(q.S) : size(q.M[1])
    γ = _pad(ntuple(µ -> _pad(ntuple(ν -> st.model.γ[µ][ν], N), MAXB, 0.0), N), MAXB, _Z4)
    bases = (hold(q.Vx), hold(q.Wx), hold(q.ex), hold(q.Vy), hold(q.Wy), hold(q.ey), hold(q.Vz), hold(q.Wz), hold(q.ez))
    if q isa TDGL.QPExchange
        m = (-q.s[2], q.s[1])
        return (Int32(1), Int32(1), Int32(q.ps.j0 - 1), Int32(nj), Int32(q.ps.k0 - 1), Int32(nk), q.σn,
                _pad(q.s, MAXB, 0.0), _pad((m[1] * st.dt, m[2] * st.dt), MAXB, 0.0), _pad((q.λ0,), MAXB, 0.0),
                _Z44, _Z44, γ, bases...)
    end
    P = _pad(ntuple(µ -> _pad(ntuple(k -> q.P[µ, k], N), MAXB, 0.0), N), MAXB, _Z4)
    Pinv = _pad(ntuple(k -> _pad(ntuple(µ -> q.Pinv[k, µ], N), MAXB, 0.0), N), MAXB, _Z4)
    (Int32(2), Int32(N), Int32(q.ps.j0 - 1), Int32(nj), Int32(q.ps.k0 - 1), Int32(nk), q.σn, _pad(q.s, MAXB, 0.0), _Z4,
     _pad(Tuple(q.λ), MAXB, 0.0), P, Pinv, γ, bases...)
end

# the OpenBC of st as the description's open fields (openbc.jl:109-140; R = 8, the table OpenBC uses)
function _open_fields(st::Stepper, hold)
    ob = st.open
    ob === nothing && return (Int32(0), Int32(0), 0.0, 0.0, _null(), _null())
    gr = st.grid; κ2 = st.model.κ^2
    (Int32(ob.every), Int32(8), gr.dx^3 / κ2, κ2, hold(TDGL.lattice_green(8)), hold(ob.Jbar))
end

_lib() = (Libdl.dlopen(libpath()); libpath())
function _check(rc, what)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:tdgl_last_error, libpath()), Cstring, ()))
    error("$what failed ($rc): $msg")
end

"""
    GPUStepper(st) -> GPUStepper

The device copy of the stepper `st` (GPUStepper(st, backend), TDGLGPU.jl:1299-1316): coefficients,
masks, work arrays and one run's state.  Errors when `st` has an option the HIP stepper does not run.
"""
mutable struct GPUStepper{N}
    st::Stepper{N}
    h::Ptr{Cvoid}
end
function GPUStepper(st::Stepper{N}) where N
    ok, why = supported(st)
    ok || error("TDGLHIP does not run $why yet (the CPU step does: TDGL.step!)")
    _lib()
    d, keep = _desc(st)
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve keep begin
        _check(ccall((:tdgl_create, libpath()), Cint, (Ref{DescC}, Ref{Ptr{Cvoid}}), d, out), "tdgl_create")
    end
    gs = GPUStepper{N}(st, out[])
    finalizer(g -> g.h == C_NULL || (ccall((:tdgl_destroy, libpath()), Cint, (Ptr{Cvoid},), g.h); g.h = C_NULL), gs)
    gs
end

"""The state of a run on the device (held by its GPUStepper's handle)."""
struct GState{N}
    gs::GPUStepper{N}
end

_ptr(a) = Ptr{Float64}(pointer(a))
_statec(s::State{N}) where N =
    StateC(_pad(ntuple(µ -> _ptr(s.G[µ]), N), MAXB, Ptr{Float64}(0)), _pad(ntuple(µ -> _ptr(s.F[µ]), N), MAXB, Ptr{Float64}(0)),
           _ptr(s.θx), _ptr(s.θy), _ptr(s.θz), _ptr(s.φx), _ptr(s.φy), _ptr(s.φz),
           _ptr(s.Ux), _ptr(s.Uy), _ptr(s.Uz), _ptr(s.Vx), _ptr(s.Vy), _ptr(s.Vz),
           _ptr(s.Jx), _ptr(s.Jy), _ptr(s.Jz), _ptr(s.hx), _ptr(s.hy), _ptr(s.hz))

"""device_state(gs, s) (TDGLGPU.jl:64-66): the state `s` on the device."""
function device_state(gs::GPUStepper{N}, s::State{N}) where N
    g = GState{N}(gs)
    device_state!(g, s)
end
"""device_state!(g, s) (TDGLGPU.jl:69-74)."""
function device_state!(g::GState{N}, s::State{N}) where N
    GC.@preserve s _check(ccall((:tdgl_upload_state, libpath()), Cint, (Ptr{Cvoid}, Ref{StateC}), g.gs.h, _statec(s)),
                          "tdgl_upload_state")
    g
end
"""host_state!(s, g): every field of the device state into `s`."""
function host_state!(s::State{N}, g::GState{N}) where N
    GC.@preserve s _check(ccall((:tdgl_download_state, libpath()), Cint, (Ptr{Cvoid}, Ref{StateC}), g.gs.h, _statec(s)),
                          "tdgl_download_state")
    s
end
"""gpu_step!(g, gs; H, HJ): one step (step!, gtdgl/src/step.jl:13-35) on the device."""
gpu_step!(g::GState, gs::GPUStepper; H::Float64, HJ::Float64) = gpu_steps!(g, gs, 1; H=H, HJ=HJ)
"""gpu_steps!(g, gs, n; H, HJ) (TDGLGPU.jl:2065-2071): n steps, synchronous."""
function gpu_steps!(g::GState, gs::GPUStepper, n::Integer; H::Float64, HJ::Float64)
    _check(ccall((:tdgl_step, libpath()), Cint, (Ptr{Cvoid}, Int32, Float64, Float64), gs.h, Int32(n), H, HJ), "tdgl_step")
    g
end

# the PCG history of the Stepper to and from the device (_pcg_up!/_pcg_down!, TDGLGPU.jl:217-224)
function _pcg_up!(gs::GPUStepper, S::PCGSolver)
    f = S.fast
    xp = f === nothing ? S.x : f.xprev
    p = PcgC(ntuple(c -> _ptr(S.x[c]), 3), ntuple(c -> _ptr(xp[c]), 3), S.warm[1], f === nothing ? 0 : f.have2[1],
             S.iters[1])
    GC.@preserve S _check(ccall((:tdgl_set_pcg, libpath()), Cint, (Ptr{Cvoid}, Ref{PcgC}), gs.h, p), "tdgl_set_pcg")
end
function _pcg_down!(S::PCGSolver, gs::GPUStepper)
    f = S.fast
    xp = f === nothing ? [zeros(size(S.x[c])) for c in 1:3] : f.xprev
    p = Ref(PcgC(ntuple(c -> _ptr(S.x[c]), 3), ntuple(c -> _ptr(xp[c]), 3), 0, 0, 0))
    GC.@preserve S xp _check(ccall((:tdgl_get_pcg, libpath()), Cint, (Ptr{Cvoid}, Ptr{PcgC}), gs.h, p), "tdgl_get_pcg")
    S.warm[1] = p[].warm != 0
    f === nothing || (f.have2[1] = p[].have2 != 0)
    S.iters[1] = p[].iters
    S
end

# the ramp of the open boundary to and from the device (_open_up!, TDGLGPU.jl:1214-1217; the host's OpenBC
# keeps the clock and the last two h_δ, which run_ladders restarts, protocol.jl:852, and open_extra saves)
struct OpenC
    count::Int32
    hold::NTuple{3,Ptr{Float64}}; hnew::NTuple{3,Ptr{Float64}}
end
function _open_up!(gs::GPUStepper, ob)
    o = OpenC(Int32(ob.count), ntuple(c -> _ptr(ob.old[c]), 3), ntuple(c -> _ptr(ob.new[c]), 3))
    GC.@preserve ob _check(ccall((:tdgl_set_open, libpath()), Cint, (Ptr{Cvoid}, Ref{OpenC}), gs.h, o), "tdgl_set_open")
end
function _open_down!(ob, gs::GPUStepper)
    o = Ref(OpenC(Int32(0), ntuple(c -> _ptr(ob.old[c]), 3), ntuple(c -> _ptr(ob.new[c]), 3)))
    GC.@preserve ob _check(ccall((:tdgl_get_open, libpath()), Cint, (Ptr{Cvoid}, Ptr{OpenC}), gs.h, o), "tdgl_get_open")
    ob.count = o[].count
    ob
end

# the normal-fluid potentials of the last step into the Stepper (TDGLGPU.jl:2136-2140)
struct QpC
    pot::NTuple{MAXB,Ptr{Float64}}
end
function _qp_down!(q, gs::GPUStepper)
    arrs = q isa TDGL.QPExchange ? (q.δ,) : Tuple(q.μ)
    p = QpC(_pad(ntuple(k -> _ptr(arrs[k]), length(arrs)), MAXB, Ptr{Float64}(0)))
    GC.@preserve arrs _check(ccall((:tdgl_get_qp, libpath()), Cint, (Ptr{Cvoid}, Ref{QpC}), gs.h, p), "tdgl_get_qp")
    q
end

# the device stream's position per Stepper (with its seed the generator's whole state; a checkpoint
# stores it where the CPU step stores noise_rng, protocol.jl:659)
const NOISE_POS = WeakKeyDict{Any,UInt64}()
"""noise_position(st) -> UInt64: the device noise stream's position of `st` after its last hook cycle."""
noise_position(st::Stepper) = get(NOISE_POS, st, UInt64(0))
noise_position!(st::Stepper, c::Integer) = (NOISE_POS[st] = UInt64(c); st)

# ---- the ladder's per-cycle measurements on the device (protocol.jl:341-352, 416-464) -------------
struct MeterC
    vi::Int32
    vo::Int32
    kappa2::Float64
    Min::Ptr{UInt8}
    Mou::Ptr{UInt8}
    line_ok::Ptr{UInt8}
    line_f1::Ptr{UInt8}
end

struct CycleRecordC
    V::Float64
    TA::Float64
    Tf1::Float64
    T::NTuple{MAXB,Float64}
    Tabs::NTuple{MAXB,Float64}
    dmax::Float64
    nbig::Int32
    pad::Int32
    It::Float64
    Is::Float64
    dif::NTuple{MAXB,Float64}
    amax::Float64
    diffh::Float64
end

"""
    meter_begin!(gs, st, delta_v) -> masks

The voltmeter of `ladder` (protocol.jl:341-352) on the device, and |Ψ| and hz of the device state as
the last cycle's (Cooper0, fieldOld).  Keep the returned masks alive while the meter is used.
"""
function meter_begin!(gs::GPUStepper, st::Stepper, delta_v::Integer)
    gr = st.grid; mk = st.masks
    (; x1, x2, y1, y2, z1, z2) = gr
    vi = x1 + delta_v; vo = x2 - delta_v
    Min = UInt8[mk.τx[vi,j,k] > 0.5 for j in y1:y2+1, k in z1:z2+1]
    Mou = UInt8[mk.τx[vo,j,k] > 0.5 for j in y1:y2+1, k in z1:z2+1]
    ok = UInt8[all(mk.τx[ii,j,k] > 0.5 for ii in vi:vo-1) for j in y1:y2+1, k in z1:z2+1]
    f1 = UInt8[ok[a,b] != 0 && all(mk.f[ii, y1+a-1, z1+b-1] >= 1 - 1e-12 for ii in vi:vo)
               for a in axes(ok, 1), b in axes(ok, 2)]
    m = Ref(MeterC(vi, vo, st.model.κ^2, pointer(Min), pointer(Mou), pointer(ok), pointer(f1)))
    GC.@preserve Min Mou ok f1 _check(ccall((:tdgl_meter_begin, libpath()), Cint, (Ptr{Cvoid}, Ptr{MeterC}),
                                            gs.h, m), "tdgl_meter_begin")
    (Min, Mou, ok, f1)
end

"""noise_counter(gs) -> UInt64: the noise stream's position (steps taken with noise); with the seed it is the
whole generator state, what a checkpoint stores in place of `noise_rng` (protocol.jl:659)."""
function noise_counter(gs::GPUStepper)
    c = Ref{UInt64}(0)
    _check(ccall((:tdgl_noise_get, libpath()), Cint, (Ptr{Cvoid}, Ref{UInt64}), gs.h, c), "tdgl_noise_get")
    c[]
end
noise_counter!(gs::GPUStepper, c::Integer) =
    _check(ccall((:tdgl_noise_set, libpath()), Cint, (Ptr{Cvoid}, UInt64), gs.h, UInt64(c)), "tdgl_noise_set")

"""noise_draws!(gs, draws, nsteps): the normals of the next nsteps steps (HOST_DRAWS[]), in the layout of
tdgl_noise_draws; `reference_draws` makes them from the Stepper's own Xoshiro stream."""
noise_draws!(gs::GPUStepper, draws::Vector{Float64}, nsteps::Integer) =
    _check(ccall((:tdgl_noise_draws, libpath()), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}), gs.h, Int32(nsteps), draws),
           "tdgl_noise_draws")

"""
    reference_draws(st, n) -> Vector{Float64}

The draws of n steps from `st.noise.rng`, taken in the reference's order (noise_kick! / noise_draw!, then
field_kick!: noise.jl:47-52, 108-111, 188-196) and laid out as tdgl_noise_draws takes them: per step the bands'
complex node arrays, then with noise_field the x, y, z link arrays.  The stream advances as the CPU step's would.
"""
function reference_draws(st::Stepper{N}, n::Integer) where N
    gr = st.grid; (; Nx, Ny, Nz, x1, x2, y1, y2, z1, z2) = gr
    rng = st.noise.rng; f = st.masks.f; hard = st.hard
    nn = (Nx + 1) * (Ny + 1) * (Nz + 1)
    shp = ((Nx, Ny + 1, Nz + 1), (Nx + 1, Ny, Nz + 1), (Nx + 1, Ny + 1, Nz))
    nl = st.noise.field ? sum(prod, shp) : 0
    out = zeros(Float64, n * (2N * nn + nl))
    p = 0
    for _ in 1:n
        for µ in 1:N
            X = zeros(ComplexF64, Nx + 1, Ny + 1, Nz + 1)
            for k in z1+1:z2, j in y1+1:y2, i in x1+1:x2
                (hard && f[i,j,k] == 0.0) && continue
                X[i,j,k] = complex(randn(rng), randn(rng))
            end
            out[p+1:p+2nn] .= reinterpret(Float64, vec(X)); p += 2nn
        end
        if st.noise.field
            rs = ((1:Nx, 2:Ny, 2:Nz), (2:Nx, 1:Ny, 2:Nz), (2:Nx, 2:Ny, 1:Nz))
            for c in 1:3
                L = zeros(shp[c])
                for k in rs[c][3], j in rs[c][2], i in rs[c][1]; L[i,j,k] = randn(rng); end
                out[p+1:p+length(L)] .= vec(L); p += length(L)
            end
        end
    end
    out
end

"""meter_cycle(gs, first) -> CycleRecordC: the record of the cycle that just ended; `first` at m == 1."""
function meter_cycle(gs::GPUStepper, first::Bool)
    r = Ref{CycleRecordC}()
    _check(ccall((:tdgl_meter_cycle, libpath()), Cint, (Ptr{Cvoid}, Int32, Ptr{CycleRecordC}), gs.h, Int32(first), r),
           "tdgl_meter_cycle")
    r[]
end

# ---- the ladder hook (TDGLGPU.jl:2076-2144) -----------------------------------------------------
const _LAST = Ref{Any}(nothing)
const _DEV = Ref{Any}(nothing)
const _FALLBACK = Ref(:cpu)
const _COUNTS = Dict(:bc => 0, :bs => 0, :cc => 0, :cs => 0)

"""execution_stats() -> (backend, backend_cycles, backend_steps, cpu_fallback_cycles, cpu_fallback_steps)"""
execution_stats() = (backend="HIP", backend_cycles=_COUNTS[:bc], backend_steps=_COUNTS[:bs],
                     cpu_fallback_cycles=_COUNTS[:cc], cpu_fallback_steps=_COUNTS[:cs])

"""
    enable!(; fallback=:cpu)

Every ladder of TDGL steps its cycles on the HIP stepper (TDGL.STEPS, gtdgl/src/protocol.jl:301-316); an
unsupported Stepper runs on the CPU with a warning, or throws with `fallback=:error`.
"""
function enable!(; fallback::Symbol=:cpu)
    fallback in (:cpu, :error) || throw(ArgumentError("fallback must be :cpu or :error"))
    _LAST[] = nothing; _DEV[] = nothing; _FALLBACK[] = fallback
    for k in keys(_COUNTS); _COUNTS[k] = 0; end
    TDGL.STEPS[] = (s, st, n; H, HJ) -> _hip_cycle!(s, st, n; H=H, HJ=HJ)
    nothing
end
disable!() = (TDGL.STEPS[] = nothing; _LAST[] = nothing; _DEV[] = nothing; nothing)

function _hip_cycle!(s::State, st::Stepper, n::Int; H::Float64, HJ::Float64)
    if _LAST[] === nothing || _LAST[][1] !== st
        ok, why = supported(st)
        if !ok
            _FALLBACK[] === :error && throw(ArgumentError("TDGLHIP: $why is unsupported; CPU fallback is disabled"))
            @warn "TDGLHIP: $why is unsupported; this stepper runs on the CPU"
        end
        _LAST[] = (st, ok ? GPUStepper(st) : nothing)
        _DEV[] = nothing
    end
    gs = _LAST[][2]
    if gs === nothing
        for _ in 1:n; step!(s, st; H=H, HJ=HJ); end
        _COUNTS[:cc] += 1; _COUNTS[:cs] += n
        return s
    end
    _DEV[] === nothing || _DEV[][1] !== s ? (_DEV[] = (s, device_state(gs, s))) : device_state!(_DEV[][2], s)
    g = _DEV[][2]
    # the Stepper's own state to the device, and back after the steps, as the CPU step leaves it
    # (TDGLGPU.jl:2130-2141): PCG history, the open boundary's clock and ramp, the noise position
    st.imp isa PCGSolver && _pcg_up!(gs, st.imp)
    st.open === nothing || _open_up!(gs, st.open)
    # with noise: the reference's own stream (HOST_DRAWS[]), or the device's stream at its position
    st.noise !== nothing && HOST_DRAWS[] && noise_draws!(gs, reference_draws(st, n), n)
    st.noise !== nothing && !HOST_DRAWS[] && haskey(NOISE_POS, st) && noise_counter!(gs, NOISE_POS[st])
    gpu_steps!(g, gs, n; H=H, HJ=HJ)
    st.imp isa PCGSolver && _pcg_down!(st.imp, gs)
    st.open === nothing || _open_down!(st.open, gs)
    st.qpx === nothing || _qp_down!(st.qpx, gs)
    st.noise !== nothing && !HOST_DRAWS[] && (NOISE_POS[st] = noise_counter(gs))
    host_state!(s, g)
    _COUNTS[:bc] += 1; _COUNTS[:bs] += n
    s
end

end # module
