"""
    PICGolf

Drop-in operators for the electrostatic PIC scripts of
jwscook/ParticleInCellCodeGolf.jl, computed on an AMD MI355X by
`libpicgolf.so` (HIP kernels for gfx950 behind the C ABI in `include/pic.h`).

The scripts' closures map onto these functions:

| script expression                                   | PICGolf                          |
|-----------------------------------------------------|----------------------------------|
| `ρ(x, y)` (src/GaussianFixedPointQuiet.jl:7)         | `deposit!(r, ctx, x, y)`         |
| `n[f(j)] += w` (src/NGPFourier.jl:5)                 | `deposit!(n, ctx, x; shape=:ngp)`|
| `real(ifft(fft(ρ)./ik))` (src/GaussianFixedPointQuiet.jl:9) | `fieldsolve!(E, ctx, ρ)` |
| `u()` (src/NGPFourier.jl:2)                          | `drift!(ctx, x, v, dt)`          |
| `u(); v += E[f.(x)]*dt` (src/NGPFourier.jl:6)        | `push!(ctx, x, v, E, dt)`        |
| `x .= X .+ (v .+ V)/2*dt` (src/GaussianFixedPointQuiet.jl:9) | `midpoint_position!(x, ctx, X, V, v, dt)` |
| `v[j] = V[j] + sum(...)*dt` (src/GaussianFixedPointQuiet.jl:10) | `push!(ctx, X, V, x, v, E, dt)` |
| `x .= mod.(x, 1)` (src/GaussianFixedPointQuiet.jl:10) | `wrap!(ctx, x)`                  |
| `isapprox(F, E, rtol=l)` (src/GaussianFixedPointQuiet.jl:8) | `isapprox(ctx, F, E; rtol)` |
| `D[t,1:4]` (src/GaussianFixedPointQuiet.jl:11,13)    | `diagnostics(ctx, E, v)`         |
| the whole `@gif for t ∈ 1:T` loop                    | `upload!`, `run!`, `download`    |
| `v[j]=V[j]+sum(...)*dt/6` etc. (src/GaussianFixedPointQuietSimpson13.jl:12,14,16) | `simpson_kick!(v, ctx, a, b, v_in, E, coef)` |

The Simpson-1/3 scripts (src/GaussianFixedPointQuietSimpson13.jl,
src/AreaFixedPointQuietSimpson13.jl) run whole with
`PicConfig(...; scheme=:simpson, shape=:gauss7 or :area)`; `download` returns
their `E[end,:]`.

Every call copies its arrays to the device and back and is synchronous; a
nonzero status from the library becomes an `error` carrying the library's
message.  The FP64 operators are bit-identical to the CPU restatement shipped
with the library (oracle/), which follows the scripts expression by
expression.
"""
module PICGolf

export Context, PicConfig, deposit!, fieldsolve!, drift!, midpoint_position!, wrap!,
       diagnostics, upload!, download, run!, run_async!, sync!, quiet_start, autotune!, set_tiling!, tiling, PicTuning,
       simpson_kick!, Pic1D2VConfig, Context1D2V, fields,
       Pic2DConfig, Context2D, gather, boris!,
       Pic2D3VSpecies, Pic2D3VConfig, Context2D3V, reset!, field, warmup!, loop!, lg_field, lg_field!, lg_bins,
       lg_grid, lg_grid!, lg_iterations,
       PicEM1DConfig, ContextEM1D, explicit_fields!, explicit_push

const libpicgolf = get(ENV, "PICGOLF_LIB", normpath(joinpath(@__DIR__, "..", "lib", "libpicgolf.so")))

const SHAPES = Dict(:ngp => Int32(0), :area => Int32(1), :gauss6 => Int32(6), :gauss7 => Int32(7))
const SCHEMES = Dict(:leapfrog => Int32(0), :midpoint => Int32(1), :simpson => Int32(2))
const PRECISIONS = Dict(:fp64 => Int32(0), :mixed => Int32(1))

"Mirror of `struct pic_config` in include/pic.h."
struct PicConfig
    N::Int32
    shape::Int32
    scheme::Int32
    precision::Int32
    P::Int64
    W::Float64
    w::Float64
    dt::Float64
    rtol::Float64
    max_sweeps::Int32
    device::Int32
end

"""
    PicConfig(; N, P, W, dt, w=W/P*N, shape=:gauss7, scheme=:midpoint,
              precision=:fp64, rtol=1e-8, max_sweeps=10, device=0)

Keyword arguments are the scripts' literals (src/GaussianFixedPointQuiet.jl:1-5).
"""
PicConfig(; N, P, W, dt, w=W / P * N, shape=:gauss7, scheme=:midpoint, precision=:fp64,
          rtol=1e-8, max_sweeps=10, device=0) =
    PicConfig(Int32(N), SHAPES[shape], SCHEMES[scheme], PRECISIONS[precision], Int64(P),
              Float64(W), Float64(w), Float64(dt), Float64(rtol), Int32(max_sweeps), Int32(device))

mutable struct Context
    ptr::Ptr{Cvoid}
    cfgs::Vector{PicConfig}
    pending::Any        # (D, S) of a run_async! run, kept reachable until sync!
end

lasterror(p::Ptr{Cvoid}) = unsafe_string(ccall((:pic_last_error, libpicgolf), Cstring, (Ptr{Cvoid},), p))

function check(ctx::Context, rc::Cint, what::AbstractString)
    rc == 0 || error(what, ": ", lasterror(ctx.ptr), " (code ", rc, ")")
    nothing
end

"""
    Context(cfg::PicConfig) / Context(cfgs::Vector{PicConfig})

One device context holding one instance or a batch of independent instances.
"""
function Context(cfgs::Vector{PicConfig})
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:pic_create_batch, libpicgolf), Cint, (Ref{Ptr{Cvoid}}, Ptr{PicConfig}, Int32),
               h, cfgs, length(cfgs))
    rc == 0 || error("pic_create_batch: ", lasterror(C_NULL), " (code ", rc, ")")
    ctx = Context(h[], cfgs, nothing)
    finalizer(destroy!, ctx)
    ctx
end
Context(cfg::PicConfig) = Context([cfg])

function destroy!(ctx::Context)
    if ctx.ptr != C_NULL
        ccall((:pic_destroy, libpicgolf), Cint, (Ptr{Cvoid},), ctx.ptr)
        ctx.ptr = C_NULL
    end
    nothing
end

Ncells(ctx::Context) = Int(ctx.cfgs[1].N)

"""
    deposit!(ρ, ctx, x, y=x; w, shape)

`r .*= 0; for l ∈ d.((x.+y)./2); for k ∈ l; r[k[1]] += k[2]*w; end; end`
(src/GaussianFixedPointQuiet.jl:7), or with `shape=:ngp` the histogram
`n[f(j)] += w` of src/NGPFourier.jl:5.  Per-cell sums are exact (integer
fixed point) and rounded once.
"""
function deposit!(ρ::Vector{Float64}, ctx::Context, x::Vector{Float64}, y::Union{Vector{Float64},Nothing}=nothing;
                  w::Real=ctx.cfgs[1].w, shape::Union{Symbol,Integer}=Int(ctx.cfgs[1].shape))
    length(ρ) == Ncells(ctx) || throw(DimensionMismatch("ρ must have N = $(Ncells(ctx)) cells"))
    y === nothing || length(y) == length(x) || throw(DimensionMismatch("x and y differ in length"))
    s = shape isa Symbol ? SHAPES[shape] : Int32(shape)
    GC.@preserve ρ x y begin
        yp = y === nothing || y === x ? C_NULL : pointer(y)
        rc = ccall((:pic_deposit, libpicgolf), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Int32, Ptr{Float64}),
                   ctx.ptr, x, yp, length(x), w, s, ρ)
    end
    check(ctx, rc, "pic_deposit")
    ρ
end

"""
    fieldsolve!(E, ctx, ρ)

`E .= real.(ifft((ξ .= fft(ρ)./ik; ξ[1] *= 0; ξ)))` (src/GaussianFixedPointQuiet.jl:9).
"""
function fieldsolve!(E::Vector{Float64}, ctx::Context, ρ::Vector{Float64})
    length(E) == length(ρ) == Ncells(ctx) || throw(DimensionMismatch("E and ρ must have N cells"))
    rc = GC.@preserve E ρ ccall((:pic_fieldsolve, libpicgolf), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                                ctx.ptr, ρ, E)
    check(ctx, rc, "pic_fieldsolve")
    E
end

"`u() = (x .= mod.(x .+ v/2*dt, 1))` (src/NGPFourier.jl:2)."
function drift!(ctx::Context, x::Vector{Float64}, v::Vector{Float64}, dt::Real)
    rc = GC.@preserve x v ccall((:pic_drift_half, libpicgolf), Cint,
                                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64), ctx.ptr, x, v, length(x), dt)
    check(ctx, rc, "pic_drift_half")
    x
end

"""
    push!(ctx, x, v, E, dt; shape)

Leapfrog second half: `u(); v += E[f.(x)]*dt` (src/NGPFourier.jl:6) for NGP,
`u(); v[j] += sum(k->E[k[1]]*k[2], d(x[j]))*dt` (src/Gaussian.jl:9-10) for the
Gaussian shape.
"""
function Base.push!(ctx::Context, x::Vector{Float64}, v::Vector{Float64}, E::Vector{Float64}, dt::Real;
                    shape::Union{Symbol,Integer}=Int(ctx.cfgs[1].shape))
    s = shape isa Symbol ? SHAPES[shape] : Int32(shape)
    rc = GC.@preserve x v E ccall((:pic_push_leapfrog, libpicgolf), Cint,
                                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Int32),
                                  ctx.ptr, x, v, E, length(x), dt, s)
    check(ctx, rc, "pic_push_leapfrog")
    (x, v)
end

"""
    push!(ctx, X, V, x, v, E, dt; shape)

Midpoint gather-kick: `v[j] = V[j] + sum(k->E[k[1]]*k[2], d((x[j]+X[j])/2))*dt`
(src/GaussianFixedPointQuiet.jl:10).
"""
function Base.push!(ctx::Context, X::Vector{Float64}, V::Vector{Float64}, x::Vector{Float64},
                    v::Vector{Float64}, E::Vector{Float64}, dt::Real;
                    shape::Union{Symbol,Integer}=Int(ctx.cfgs[1].shape))
    s = shape isa Symbol ? SHAPES[shape] : Int32(shape)
    rc = GC.@preserve X V x v E ccall((:pic_push_midpoint, libpicgolf), Cint,
                                      (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                                       Ptr{Float64}, Int64, Float64, Int32),
                                      ctx.ptr, X, V, x, v, E, length(X), dt, s)
    check(ctx, rc, "pic_push_midpoint")
    v
end

"""
    simpson_kick!(v_out, ctx, a, b, v_in, E, coef; shape)

One stage kick of the Simpson-1/3 sweep:
`v_out[j] = v_in[j] + sum(k->E[k[1]]*k[2], d((a[j]+b[j])/2))*coef/6`
(src/GaussianFixedPointQuietSimpson13.jl:12,14,16 with `(a, b, coef)` =
`(X, X, dt)`, `(X, x, 4dt)`, `(x, x, dt)`; src/AreaFixedPointQuietSimpson13.jl:11,13,15
with `shape=:area`).
"""
function simpson_kick!(v_out::Vector{Float64}, ctx::Context, a::Vector{Float64}, b::Vector{Float64},
                       v_in::Vector{Float64}, E::Vector{Float64}, coef::Real;
                       shape::Union{Symbol,Integer}=Int(ctx.cfgs[1].shape))
    s = shape isa Symbol ? SHAPES[shape] : Int32(shape)
    rc = GC.@preserve v_out a b v_in E ccall((:pic_simpson_kick, libpicgolf), Cint,
                                             (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                                              Ptr{Float64}, Int64, Float64, Int32),
                                             ctx.ptr, a, b, v_in, v_out, E, length(a), coef, s)
    check(ctx, rc, "pic_simpson_kick")
    v_out
end

"`x .= X .+ (v .+ V)/2*dt` (src/GaussianFixedPointQuiet.jl:9)."
function midpoint_position!(x::Vector{Float64}, ctx::Context, X::Vector{Float64}, V::Vector{Float64},
                            v::Vector{Float64}, dt::Real)
    rc = GC.@preserve x X V v ccall((:pic_midpoint_position, libpicgolf), Cint,
                                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64),
                                    ctx.ptr, X, V, v, x, length(X), dt)
    check(ctx, rc, "pic_midpoint_position")
    x
end

"`x .= mod.(x, 1)` (src/GaussianFixedPointQuiet.jl:10)."
function wrap!(ctx::Context, x::Vector{Float64})
    rc = GC.@preserve x ccall((:pic_wrap, libpicgolf), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64), ctx.ptr, x, length(x))
    check(ctx, rc, "pic_wrap")
    x
end

"`isapprox(F, E, rtol=rtol, atol=atol)` (src/GaussianFixedPointQuiet.jl:8), computed on the GPU."
function Base.isapprox(ctx::Context, F::Vector{Float64}, E::Vector{Float64}; rtol::Real, atol::Real=0)
    out = Ref{Int32}(0)
    rc = GC.@preserve F E ccall((:pic_isapprox, libpicgolf), Cint,
                                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int32, Float64, Float64, Ref{Int32}),
                                ctx.ptr, F, E, length(F), rtol, atol, out)
    check(ctx, rc, "pic_isapprox")
    out[] != 0
end

"`D[t,1:4]` of src/GaussianFixedPointQuiet.jl:11,13 for one state."
function diagnostics(ctx::Context, E::Vector{Float64}, v::Vector{Float64}; W::Real=ctx.cfgs[1].W)
    out = zeros(4)
    rc = GC.@preserve E v out ccall((:pic_diagnostics, libpicgolf), Cint,
                                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Ptr{Float64}),
                                    ctx.ptr, E, v, length(v), W, out)
    check(ctx, rc, "pic_diagnostics")
    out
end

"Upload (x, v) and the field carried between steps (zeros as in the scripts) for instance `inst` (1-based)."
function upload!(ctx::Context, x::Vector{Float64}, v::Vector{Float64}, E::Union{Vector{Float64},Nothing}=nothing;
                 inst::Integer=1)
    GC.@preserve x v E begin
        Ep = E === nothing ? C_NULL : pointer(E)
        rc = ccall((:pic_upload, libpicgolf), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   ctx.ptr, inst - 1, x, v, Ep)
    end
    check(ctx, rc, "pic_upload")
    ctx
end

"Download (x, v, E) of instance `inst` (1-based)."
function download(ctx::Context; inst::Integer=1)
    c = ctx.cfgs[inst]
    x, v, E = zeros(c.P), zeros(c.P), zeros(c.N)
    rc = GC.@preserve x v E ccall((:pic_download, libpicgolf), Cint,
                                  (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                  ctx.ptr, inst - 1, x, v, E)
    check(ctx, rc, "pic_download")
    (x, v, E)
end

"""
    run!(ctx, nsteps) -> (D, sweeps)

Advance every instance `nsteps` steps on the device.  `D[t, :, i]` is the
script's `D[t, 1:4]` for instance i; `sweeps[t, i]` the Picard sweeps taken.
"""
function run!(ctx::Context, nsteps::Integer)
    n = length(ctx.cfgs)
    D = zeros(4, nsteps, n)          # C row-major [inst][step][4]
    S = zeros(Int32, nsteps, n)
    rc = GC.@preserve D S ccall((:pic_run, libpicgolf), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Int32}),
                                ctx.ptr, nsteps, D, S)
    check(ctx, rc, "pic_run")
    (permutedims(D, (2, 1, 3)), S)
end

"""
    run_async!(ctx, nsteps)

Starts `nsteps` steps on a thread of the library (`pic_run_async`) and returns
at once; `sync!(ctx)` waits and returns what `run!` would.  Other calls on
the context fail until then.
"""
function run_async!(ctx::Context, nsteps::Integer)
    n = length(ctx.cfgs)
    D = zeros(4, nsteps, n)
    S = zeros(Int32, nsteps, n)
    rc = ccall((:pic_run_async, libpicgolf), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Int32}),
               ctx.ptr, nsteps, D, S)
    check(ctx, rc, "pic_run_async")
    ctx.pending = (D, S)
    nothing
end

function sync!(ctx::Context)
    rc = ccall((:pic_sync, libpicgolf), Cint, (Ptr{Cvoid},), ctx.ptr)
    p, ctx.pending = ctx.pending, nothing
    check(ctx, rc, "pic_sync")
    p === nothing ? nothing : (permutedims(p[1], (2, 1, 3)), p[2])
end

"Mirror of `struct pic_tuning` in include/pic.h."
struct PicTuning
    tile_particles::Int64
    block::Int32
    n_tiles::Int32
    candidates::Int32
    ms_per_step::Float64
    entry_ms_per_step::Float64
end

"""
    autotune!(ctx; trial_steps=2) -> PicTuning

Times `trial_steps` steps of the current state under each candidate particle
tiling of the push launches, keeps the fastest and restores the state.  Runs
give the same results under any tiling.
"""
function autotune!(ctx::Context; trial_steps::Integer=2)
    r = Ref{PicTuning}()
    rc = ccall((:pic_autotune, libpicgolf), Cint, (Ptr{Cvoid}, Int32, Ref{PicTuning}), ctx.ptr, trial_steps, r)
    check(ctx, rc, "pic_autotune")
    r[]
end

"`set_tiling!(ctx, tile_particles, block)`: the push launch layout (results do not depend on it)."
function set_tiling!(ctx::Context, tile_particles::Integer, block::Integer=1024)
    rc = ccall((:pic_set_tiling, libpicgolf), Cint, (Ptr{Cvoid}, Int64, Int32), ctx.ptr, tile_particles, block)
    check(ctx, rc, "pic_set_tiling")
    tiling(ctx)
end

"`tiling(ctx) -> (tile_particles, block, n_tiles)`."
function tiling(ctx::Context)
    t, b, n = Ref{Int64}(0), Ref{Int32}(0), Ref{Int32}(0)
    rc = ccall((:pic_get_tiling, libpicgolf), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int32}, Ref{Int32}),
               ctx.ptr, t, b, n)
    check(ctx, rc, "pic_get_tiling")
    (t[], b[], n[])
end

"`x = (bitreverse.(0:P-1) .+ 2.0^63)/2.0^64`, `v = ∓1` by half (src/GaussianFixedPointQuiet.jl:2-3)."
quiet_start(P::Integer) = ((bitreverse.(0:P-1) .+ 2.0^63) / 2.0^64, collect(1:P .> P / 2) .* 2 .- 1.0)

# ---------------------------------------------------------------------------
# Electrostatic 2D3V (src/Electrostatic2D3V.jl), C ABI in include/pic2d.h.
# Grids are Matrix{Float64}(undef, NX, NY): the script's own layout.
#
# | script expression                                   | PICGolf                               |
# |-----------------------------------------------------|---------------------------------------|
# | `eval(Ex, Ey, x[i], y[i])` (:94-103)                 | `gather(ctx, Ex, Ey, x, y)`           |
# | `boris(vx, vy, vz, Ex, Ey, dt)` (:32-41)             | `boris!(ctx, vx, vy, vz, ex, ey, dt)` |
# | particle loop without the deposit (:128-134)         | `push!(ctx, x, y, vx, vy, vz, Ex, Ey)`|
# | `depositcharge!` + `sum(ns, dims=3)` (:105-109,139-141) | `deposit!(ρ, ctx, x, y)`          |
# | field invert + field solve (:142-157)                | `fieldsolve!(Ex, Ey, ctx, ρ)`         |
# | `K[ti, 1:5]` (:166-170)                              | `diagnostics(ctx, Ex, Ey, vx, vy)`    |
# | the `for t in 1:T` loop (:120-176)                   | `upload!`, `run!`, `download`         |
# ---------------------------------------------------------------------------

const PAIRINGS = Dict(:reference => Int32(0), :textbook => Int32(1))

"Mirror of `struct pic2d_config` in include/pic2d.h."
struct Pic2DConfig
    NX::Int32
    NY::Int32
    P::Int64
    precision::Int32
    pairing::Int32
    dt::Float64
    w::Float64
    B0::Float64
    ns::Int32
    device::Int32
    sort_every::Int32
end

"""
    Pic2DConfig(; NX, NY, P, dt, w, B0, ns=2, precision=:fp64, pairing=:reference, device=0, sort_every=0)

`w = n0/P/(Δx*Δy)`, `B0` along x and `NS` as in src/Electrostatic2D3V.jl:23-25.
"""
Pic2DConfig(; NX, NY, P, dt, w, B0, ns=2, precision=:fp64, pairing=:reference, device=0, sort_every=0) =
    Pic2DConfig(Int32(NX), Int32(NY), Int64(P), PRECISIONS[precision], PAIRINGS[pairing], Float64(dt),
                Float64(w), Float64(B0), Int32(ns), Int32(device), Int32(sort_every))

mutable struct Context2D
    ptr::Ptr{Cvoid}
    cfg::Pic2DConfig
    steps::Int64
end

lasterror2d(p::Ptr{Cvoid}) = unsafe_string(ccall((:pic2d_last_error, libpicgolf), Cstring, (Ptr{Cvoid},), p))

function check(ctx::Context2D, rc::Cint, what::AbstractString)
    rc == 0 || error(what, ": ", lasterror2d(ctx.ptr), " (code ", rc, ")")
    nothing
end

function Context2D(cfg::Pic2DConfig)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:pic2d_create, libpicgolf), Cint, (Ref{Ptr{Cvoid}}, Ref{Pic2DConfig}), h, cfg)
    rc == 0 || error("pic2d_create: ", lasterror2d(C_NULL), " (code ", rc, ")")
    ctx = Context2D(h[], cfg, 0)
    finalizer(destroy!, ctx)
    ctx
end

function destroy!(ctx::Context2D)
    if ctx.ptr != C_NULL
        ccall((:pic2d_destroy, libpicgolf), Cint, (Ptr{Cvoid},), ctx.ptr)
        ctx.ptr = C_NULL
    end
    nothing
end

grid(ctx::Context2D) = zeros(ctx.cfg.NX, ctx.cfg.NY)

"`eval(Ex, Ey, x[p], y[p])` for every particle (src/Electrostatic2D3V.jl:94-103)."
function gather(ctx::Context2D, Ex::Matrix{Float64}, Ey::Matrix{Float64}, x::Vector{Float64}, y::Vector{Float64})
    ex, ey = similar(x), similar(x)
    rc = GC.@preserve Ex Ey x y ex ey ccall((:pic2d_gather, libpicgolf), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}),
        ctx.ptr, Ex, Ey, x, y, length(x), ex, ey)
    check(ctx, rc, "pic2d_gather")
    (ex, ey)
end

"`vx, vy, vz = boris(vx, vy, vz, ex, ey, dt)` in place (src/Electrostatic2D3V.jl:32-41)."
function boris!(ctx::Context2D, vx::Vector{Float64}, vy::Vector{Float64}, vz::Vector{Float64},
                ex::Vector{Float64}, ey::Vector{Float64}, dt::Real=ctx.cfg.dt; B0::Real=ctx.cfg.B0)
    rc = GC.@preserve vx vy vz ex ey ccall((:pic2d_boris, libpicgolf), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Float64),
        ctx.ptr, vx, vy, vz, ex, ey, length(vx), dt, B0)
    check(ctx, rc, "pic2d_boris")
    (vx, vy, vz)
end

"Gather, Boris and `x = unimod(x + vx*dt, 1)` in place (src/Electrostatic2D3V.jl:128-134)."
function Base.push!(ctx::Context2D, x::Vector{Float64}, y::Vector{Float64}, vx::Vector{Float64},
                    vy::Vector{Float64}, vz::Vector{Float64}, Ex::Matrix{Float64}, Ey::Matrix{Float64};
                    dt::Real=ctx.cfg.dt, B0::Real=ctx.cfg.B0)
    rc = GC.@preserve x y vx vy vz Ex Ey ccall((:pic2d_push, libpicgolf), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64,
         Ptr{Float64}, Ptr{Float64}, Float64, Float64),
        ctx.ptr, x, y, vx, vy, vz, length(x), Ex, Ey, dt, B0)
    check(ctx, rc, "pic2d_push")
    (x, y, vx, vy, vz)
end

"`depositcharge!` of every particle, summed over threads (src/Electrostatic2D3V.jl:105-109,139-141)."
function deposit!(ρ::Matrix{Float64}, ctx::Context2D, x::Vector{Float64}, y::Vector{Float64}; w::Real=ctx.cfg.w)
    size(ρ) == (ctx.cfg.NX, ctx.cfg.NY) || throw(DimensionMismatch("ρ must be NX x NY"))
    rc = GC.@preserve ρ x y ccall((:pic2d_deposit, libpicgolf), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Ptr{Float64}), ctx.ptr, x, y, length(x), w, ρ)
    check(ctx, rc, "pic2d_deposit")
    ρ
end

"""
    fieldsolve!(Ex, Ey, ctx, ρ; phi=nothing)

`pfft*phi; phi[1,1] = 0; Ex = phi.*minvkk.*kx; Ey = ...; pifft*Ex; pifft*Ey`
(src/Electrostatic2D3V.jl:142-157), real parts out; `phi` receives
`real(pifft*phi)` as stored in `phis` (:173).
"""
function fieldsolve!(Ex::Matrix{Float64}, Ey::Matrix{Float64}, ctx::Context2D, ρ::Matrix{Float64};
                     phi::Union{Matrix{Float64},Nothing}=nothing)
    GC.@preserve Ex Ey ρ phi begin
        pp = phi === nothing ? C_NULL : pointer(phi)
        rc = ccall((:pic2d_fieldsolve, libpicgolf), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}), ctx.ptr, ρ, Ex, Ey, pp)
    end
    check(ctx, rc, "pic2d_fieldsolve")
    (Ex, Ey)
end

"`K[ti, 1:5]` of src/Electrostatic2D3V.jl:166-170 for one state."
function diagnostics(ctx::Context2D, Ex::Matrix{Float64}, Ey::Matrix{Float64}, vx::Vector{Float64},
                     vy::Vector{Float64}; w::Real=ctx.cfg.w)
    K = zeros(5)
    rc = GC.@preserve Ex Ey vx vy K ccall((:pic2d_diagnostics, libpicgolf), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Ptr{Float64}),
        ctx.ptr, Ex, Ey, vx, vy, length(vx), w, K)
    check(ctx, rc, "pic2d_diagnostics")
    K
end

"Upload the particles (x, y in (0, 1]) and the field the next step gathers from (zeros by default)."
function upload!(ctx::Context2D, x, y, vx, vy, vz; Ex::Union{Matrix{Float64},Nothing}=nothing,
                 Ey::Union{Matrix{Float64},Nothing}=nothing)
    GC.@preserve x y vx vy vz Ex Ey begin
        ep = Ex === nothing ? C_NULL : pointer(Ex)
        fp = Ey === nothing ? C_NULL : pointer(Ey)
        rc = ccall((:pic2d_upload, libpicgolf), Cint,
                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                    Ptr{Float64}, Ptr{Float64}), ctx.ptr, x, y, vx, vy, vz, ep, fp)
    end
    check(ctx, rc, "pic2d_upload")
    ctx.steps = 0
    ctx
end

"Download (x, y, vx, vy, vz, Ex, Ey)."
function download(ctx::Context2D)
    P = ctx.cfg.P
    out = (zeros(P), zeros(P), zeros(P), zeros(P), zeros(P), grid(ctx), grid(ctx))
    rc = GC.@preserve out ccall((:pic2d_download, libpicgolf), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}), ctx.ptr, out...)
    check(ctx, rc, "pic2d_download")
    out
end

"""
    run!(ctx::Context2D, nsteps) -> K

Advance `nsteps` steps; `K` holds one row `K[ti, 1:5]` per step t with
`t % NS == 0` (src/Electrostatic2D3V.jl:164-169).
"""
function run!(ctx::Context2D, nsteps::Integer)
    ns = ctx.cfg.ns
    rows = div(ctx.steps + nsteps, ns) - div(ctx.steps, ns)
    K = zeros(5, rows)
    rc = GC.@preserve K ccall((:pic2d_run, libpicgolf), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}),
                              ctx.ptr, nsteps, rows > 0 ? pointer(K) : C_NULL)
    check(ctx, rc, "pic2d_run")
    ctx.steps += nsteps
    permutedims(K)
end

# ---------------------------------------------------------------------------
# Magnetised 1D2V (src/NGP1D2V.jl, src/NGP1D2V2S.jl), C ABI in include/pic1d2v.h.
#
# | script expression                                      | PICGolf                                  |
# |--------------------------------------------------------|------------------------------------------|
# | `rho()` / `rho(x)` (NGP1D2V2S.jl:28-29, NGP1D2V.jl:38)  | `deposit(ctx, xs)`                        |
# | `sum(k->E[k[1]]*k[2], d(x[j]))` (:35, :42)              | `gather(ctx, E, x)`                       |
# | `boris(vx, vy, E, B, dt, q_m)` (:5-11)                   | `boris!(ctx, vx, vy, Ej; q_m)`            |
# | the `for t in 1:T` loop with `Es` and `D` (:31-57)       | `upload!`, `run!`, `fields`, `download`   |
# ---------------------------------------------------------------------------

"Mirror of `struct pic1d2v_config` in include/pic1d2v.h."
struct Pic1D2VConfig
    N::Int32
    shape::Int32
    nspecies::Int32
    bin::Int32
    P::Int64
    q::NTuple{2,Float64}
    q_m::NTuple{2,Float64}
    mass::NTuple{2,Float64}
    w::Float64
    n0::Float64
    B0::Float64
    dt::Float64
    device::Int32
    precision::Int32
end

"""
    Pic1D2VConfig(; N, P, bin, w, n0, B0, dt, q=(1.0,), q_m=(1.0,), mass=(1.0,), shape=:gauss7, device=0,
                  precision=:fp64)

One species (src/NGP1D2V.jl: `q = q_m = mass = (1.0,)`, `w = n0/P`) or two
(src/NGP1D2V2S.jl: `q = (-1.0, 1.0)`, `q_m = (-1.0, 1/M)`, `mass = (1.0, M)`,
`w = n0/2P`); `bin = T÷TO`.  `precision = :mixed` runs the loop with FP32
particle arithmetic and exact integer charge sums.
"""
function Pic1D2VConfig(; N, P, bin, w, n0, B0, dt, q=(1.0,), q_m=(1.0,), mass=(1.0,), shape=:gauss7, device=0,
                       precision=:fp64)
    ns = length(q)
    pad(t) = (Float64(t[1]), ns == 2 ? Float64(t[2]) : 0.0)
    Pic1D2VConfig(Int32(N), SHAPES[shape], Int32(ns), Int32(bin), Int64(P), pad(q), pad(q_m), pad(mass),
                  Float64(w), Float64(n0), Float64(B0), Float64(dt), Int32(device), PRECISIONS[precision])
end

mutable struct Context1D2V
    ptr::Ptr{Cvoid}
    cfg::Pic1D2VConfig
    steps::Int64
end

function check(ctx::Context1D2V, rc::Cint, what::AbstractString)
    rc == 0 || error("$what: ", unsafe_string(ccall((:pic1d2v_last_error, libpicgolf), Cstring, (Ptr{Cvoid},), ctx.ptr)))
    nothing
end

function Context1D2V(cfg::Pic1D2VConfig)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:pic1d2v_create, libpicgolf), Cint, (Ref{Ptr{Cvoid}}, Ref{Pic1D2VConfig}), h, cfg)
    rc == 0 || error("pic1d2v_create: ", unsafe_string(ccall((:pic1d2v_last_error, libpicgolf), Cstring,
                                                             (Ptr{Cvoid},), C_NULL)))
    ctx = Context1D2V(h[], cfg, 0)
    finalizer(destroy!, ctx)
    ctx
end

function destroy!(ctx::Context1D2V)
    if ctx.ptr != C_NULL
        ccall((:pic1d2v_destroy, libpicgolf), Cint, (Ptr{Cvoid},), ctx.ptr)
        ctx.ptr = C_NULL
    end
    nothing
end

"`upload!(ctx, species, x, vx, vy)` (species 1 or 2); restarts the step count and the bins."
function upload!(ctx::Context1D2V, species::Integer, x::Vector{Float64}, vx::Vector{Float64}, vy::Vector{Float64})
    rc = GC.@preserve x vx vy ccall((:pic1d2v_upload, libpicgolf), Cint,
                                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                    ctx.ptr, species - 1, x, vx, vy)
    check(ctx, rc, "pic1d2v_upload")
    ctx.steps = 0
    ctx
end

"`download(ctx, species) -> (x, vx, vy)`."
function download(ctx::Context1D2V, species::Integer)
    P = ctx.cfg.P
    x, vx, vy = zeros(P), zeros(P), zeros(P)
    rc = GC.@preserve x vx vy ccall((:pic1d2v_download, libpicgolf), Cint,
                                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                    ctx.ptr, species - 1, x, vx, vy)
    check(ctx, rc, "pic1d2v_download")
    (x, vx, vy)
end

"""
    run!(ctx::Context1D2V, nsteps) -> D

Advance `nsteps` steps; `D` holds the rows `D[ti, 1:5]` of the bins completed
in this call, divided by the bin length as the scripts' `D ./= T÷TO` does.
"""
function run!(ctx::Context1D2V, nsteps::Integer)
    b = ctx.cfg.bin
    rows = div(ctx.steps + nsteps, b) - div(ctx.steps, b)
    D = zeros(5, rows)
    rc = GC.@preserve D ccall((:pic1d2v_run, libpicgolf), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}),
                              ctx.ptr, nsteps, rows > 0 ? pointer(D) : C_NULL)
    check(ctx, rc, "pic1d2v_run")
    ctx.steps += nsteps
    permutedims(D) ./ b
end

"`fields(ctx) -> Es`: the scripts' `Es` (N x bins so far, `./= T÷TO`)."
function fields(ctx::Context1D2V)
    b = ctx.cfg.bin
    nb = cld(ctx.steps, b)
    Es = zeros(ctx.cfg.N, nb)
    if nb > 0
        rc = GC.@preserve Es ccall((:pic1d2v_fields, libpicgolf), Cint, (Ptr{Cvoid}, Ptr{Float64}, Int64),
                                   ctx.ptr, Es, nb)
        check(ctx, rc, "pic1d2v_fields")
    end
    Es ./ b
end

"`deposit(ctx, xs) -> r`: `rho()` over the species' positions (src/NGP1D2V2S.jl:28-29)."
function deposit(ctx::Context1D2V, xs::Vector{Vector{Float64}})
    x = reduce(vcat, xs)
    r = zeros(ctx.cfg.N)
    rc = GC.@preserve x r ccall((:pic1d2v_deposit, libpicgolf), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                                ctx.ptr, x, r)
    check(ctx, rc, "pic1d2v_deposit")
    r
end

"`gather(ctx, E, x) -> Ej`: `sum(k->E[k[1]]*k[2], d(x[j]))` for every j."
function gather(ctx::Context1D2V, E::Vector{Float64}, x::Vector{Float64})
    Ej = zeros(length(x))
    rc = GC.@preserve E x Ej ccall((:pic1d2v_gather, libpicgolf), Cint,
                                   (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}),
                                   ctx.ptr, E, x, length(x), Ej)
    check(ctx, rc, "pic1d2v_gather")
    Ej
end

"`boris!(ctx, vx, vy, Ej; B, dt, q_m)`: `vx[j], vy[j] = boris(vx[j], vy[j], Ej[j], B, dt, q_m)` in place."
function boris!(ctx::Context1D2V, vx::Vector{Float64}, vy::Vector{Float64}, Ej::Vector{Float64};
                B::Real=ctx.cfg.B0, dt::Real=ctx.cfg.dt, q_m::Real=1.0)
    rc = GC.@preserve vx vy Ej ccall((:pic1d2v_boris, libpicgolf), Cint,
                                     (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64,
                                      Float64, Float64), ctx.ptr, vx, vy, Ej, length(vx), B, dt, q_m)
    check(ctx, rc, "pic1d2v_boris")
    (vx, vy)
end

# ---------------------------------------------------------------------------
# Structured 2D3V module, electrostatic path (src/PIC2D3V.jl), C ABI in
# include/pic2d3v.h.
#
# | PIC2D3V                                                  | PICGolf                                   |
# |----------------------------------------------------------|-------------------------------------------|
# | `Species(...)`, `species.xyv` (:148-208)                  | `Pic2D3VSpecies`, `upload!`, `download`    |
# | `ElectrostaticField(NX, NY, Lx, Ly; dt, B0x, B0y, B0z, buffer)` (:284-292) | `Pic2D3VConfig(...)`     |
# | `ElectrostaticDiagnostics(NX, NY, NT, ntskip, ngskip)` (:95-101) | `Pic2D3VConfig(...; ntskip, ngskip)` |
# | `sort!(species, Δx, Δy)` (:210-215)                       | `sort!(ctx, s, Δx, Δy)`                    |
# | `loop!(plasma, field, to, t, _); diagnose!(d, field, plasma, t, to)` (:530-581, :1310-1330) | `run!` |
# | `d.Exs, d.Eys, d.ϕs` (:1323-1327)                          | `fields(ctx)`                              |
# | `field.Exy` (:294-297)                                     | `field(ctx)`                               |
# | `deposit!(z, shape, x, y, NX_Lx, NY_Ly, w)` (:1249-1256)   | `deposit!(ρ, ctx, shape, x, y, w)`         |
# | `(field)(shape, x, y)` (:1217-1228)                        | `gather(ctx, shape, Ex, Ey, x, y)`         |
# | `field.boris(vx, vy, vz, Ex, Ey, q_m)` (:49-54)            | `boris!(ctx, vx, vy, vz, Ex, Ey; q_m)`     |
# ---------------------------------------------------------------------------

const SHAPES2D3V = Dict(:ngp => Int32(0), :area => Int32(1), :bspline0 => Int32(2), :bspline1 => Int32(3),
                        :bspline2 => Int32(4), :bspline3 => Int32(5), :bspline4 => Int32(6), :bspline5 => Int32(7))
const FIELD_UPDATES = Dict(:accumulate => Int32(0), :replace => Int32(1))
const BORIS_VARIANTS = Dict(:reference => Int32(0), :textbook => Int32(1))
const FIELDS2D3V = Dict(:electrostatic => Int32(0), :lorenz_staggered => Int32(1), :lorenz => Int32(2),
                        :lorenz_semi_implicit => Int32(3))
const MAX_SPECIES_2D3V = 4
# the Lorenz-gauge state slots and binned fields of include/pic2d3v.h
const LG_STATE = (:ϕ⁰, :ϕ⁻, :Ax⁰, :Ay⁰, :Az⁰, :Ax⁻, :Ay⁻, :Az⁻, :ρ⁰, :Jx⁰, :Jy⁰, :Jz⁰, :Ex, :Ey, :Ez, :Bx, :By, :Bz,
                  :ρ⁻, :Jx⁻, :Jy⁻, :Jz⁻)  # the last four: LorenzGaugeSemiImplicitField only
const LG_BINS = (:Exs, :Eys, :Ezs, :Bxs, :Bys, :Bzs, :Axs, :Ays, :Azs, :ϕs, :ρs, :Jxs, :Jys, :Jzs)

"Mirror of `struct pic2d3v_species` in include/pic2d3v.h."
struct Pic2D3VSpecies
    P::Int64
    shape::Int32
    charge::Float64
    mass::Float64
    weight::Float64
end

"""
    Pic2D3VSpecies(P; shape=:bspline2, charge=1, mass=1, weight)

One `Species` (src/PIC2D3V.jl:148-156); `weight` is the reference's
`calculateweight(n0, P, Lx, Ly) = n0*Lx*Ly/P`.
"""
Pic2D3VSpecies(P::Integer; shape=:bspline2, charge=1.0, mass=1.0, weight) =
    Pic2D3VSpecies(Int64(P), SHAPES2D3V[shape], Float64(charge), Float64(mass), Float64(weight))

"Mirror of `struct pic2d3v_config` in include/pic2d3v.h."
struct Pic2D3VConfig
    NX::Int32
    NY::Int32
    Lx::Float64
    Ly::Float64
    dt::Float64
    B0::NTuple{3,Float64}
    buffer::Int32
    ntskip::Int32
    ngskip::Int32
    field_update::Int32
    boris::Int32
    nspecies::Int32
    species::NTuple{4,Pic2D3VSpecies}
    device::Int32
    field::Int32
    theta::Float64
    theta_source::Float64
    rtol::Float64
    maxiters::Int32
    precision::Int32
end

"""
    Pic2D3VConfig(species; NX, NY=NX, Lx=1, Ly=1, dt, B0x=0, B0y=0, B0z=0, buffer=3, ntskip=1, ngskip=1,
                  field_update=:accumulate, boris=:reference, device=0, field=:electrostatic, theta=1,
                  theta_source=0, rtol=sqrt(eps()), maxiters=10, precision=:fp64)

The `ElectrostaticField(NX, NY, Lx, Ly; dt, B0x, B0y, B0z, buffer)` and
`ElectrostaticDiagnostics(..., ntskip, ngskip)` of a plasma of up to four
species.  `field_update=:accumulate` is the reference's `update!`, which adds
every step's field into `Exy` without clearing it; `:replace` gathers the
step's own field.  `boris=:reference` is `ElectrostaticBoris` (rotation scaled
by q_m^2); `:textbook` is `ElectromagneticBoris`'s form with Ez = 0.
`field=:lorenz_staggered` builds `LorenzGaugeStaggeredField(NX, NY, Lx, Ly; dt,
B0x, B0y, B0z, imex=ImEx(theta), buffer)` and `LorenzGaugeDiagnostics` instead
(`theta=0` is `Explicit()`, `1` is `Implicit()`; src/2D3V.jl uses `ImEx(1)`);
`field=:lorenz` builds a `LorenzGaugeField`, which runs with `loop!` only;
`field=:lorenz_semi_implicit` a `LorenzGaugeSemiImplicitField(...; fieldimex=ImEx(theta),
sourceimex=ImEx(theta_source), buffer, rtol, maxiters)`.  `precision=:mixed`
(electrostatic field only) runs the particles in mixed precision: fixed-point
positions, FP32 hi + lo velocities, FP32 gathers and exact integer deposits.
"""
function Pic2D3VConfig(species::Vector{Pic2D3VSpecies}; NX, NY=NX, Lx=1.0, Ly=1.0, dt, B0x=0.0, B0y=0.0, B0z=0.0,
                       buffer=3, ntskip=1, ngskip=1, field_update=:accumulate, boris=:reference, device=0,
                       field=:electrostatic, theta=1.0, theta_source=0.0, rtol=sqrt(eps()), maxiters=10,
                       precision=:fp64)
    ns = length(species)
    1 <= ns <= MAX_SPECIES_2D3V || error("Pic2D3VConfig: 1 to $MAX_SPECIES_2D3V species")
    none = Pic2D3VSpecies(0, 0, 0.0, 0.0, 0.0)
    sp = ntuple(k -> k <= ns ? species[k] : none, 4)
    Pic2D3VConfig(Int32(NX), Int32(NY), Float64(Lx), Float64(Ly), Float64(dt),
                  (Float64(B0x), Float64(B0y), Float64(B0z)), Int32(buffer), Int32(ntskip), Int32(ngskip),
                  FIELD_UPDATES[field_update], BORIS_VARIANTS[boris], Int32(ns), sp, Int32(device),
                  FIELDS2D3V[field], Float64(theta), Float64(theta_source), Float64(rtol), Int32(maxiters),
                  PRECISIONS[precision])
end

mutable struct Context2D3V
    ptr::Ptr{Cvoid}
    cfg::Pic2D3VConfig
    t::Int64
end

function check(ctx::Context2D3V, rc::Cint, what::AbstractString)
    rc == 0 || error("$what: ", unsafe_string(ccall((:pic2d3v_last_error, libpicgolf), Cstring, (Ptr{Cvoid},), ctx.ptr)))
    nothing
end

function Context2D3V(cfg::Pic2D3VConfig)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:pic2d3v_create, libpicgolf), Cint, (Ref{Ptr{Cvoid}}, Ref{Pic2D3VConfig}), h, cfg)
    rc == 0 || error("pic2d3v_create: ", unsafe_string(ccall((:pic2d3v_last_error, libpicgolf), Cstring,
                                                             (Ptr{Cvoid},), C_NULL)))
    ctx = Context2D3V(h[], cfg, 0)
    finalizer(destroy!, ctx)
    ctx
end

function destroy!(ctx::Context2D3V)
    if ctx.ptr != C_NULL
        ccall((:pic2d3v_destroy, libpicgolf), Cint, (Ptr{Cvoid},), ctx.ptr)
        ctx.ptr = C_NULL
    end
    nothing
end

"`upload!(ctx, s, xyv)`: species `s`'s 5 x P matrix `xyv` (rows x, y, vx, vy, vz)."
function upload!(ctx::Context2D3V, s::Integer, xyv::Matrix{Float64})
    size(xyv) == (5, ctx.cfg.species[s].P) || error("upload!: xyv must be 5 x $(ctx.cfg.species[s].P)")
    rc = GC.@preserve xyv ccall((:pic2d3v_upload, libpicgolf), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                                ctx.ptr, s - 1, xyv)
    check(ctx, rc, "pic2d3v_upload")
    ctx
end

"`download(ctx, s) -> xyv` (5 x P)."
function download(ctx::Context2D3V, s::Integer)
    xyv = zeros(5, ctx.cfg.species[s].P)
    rc = GC.@preserve xyv ccall((:pic2d3v_download, libpicgolf), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}),
                                ctx.ptr, s - 1, xyv)
    check(ctx, rc, "pic2d3v_download")
    xyv
end

"`reset!(ctx)`: a fresh field and diagnostics (Exy = 0, t = 0)."
function reset!(ctx::Context2D3V)
    check(ctx, ccall((:pic2d3v_reset, libpicgolf), Cint, (Ptr{Cvoid},), ctx.ptr), "pic2d3v_reset")
    ctx.t = 0
    ctx
end

"""
    sort!(ctx::Context2D3V, s, Δx, Δy) -> p

`sort!(species, Δx, Δy)`: reorders species `s` by `(ceil(x/Δx), ceil(y/Δy), vx)`
(stable) and returns the permutation `species.p` (1-based).
"""
function Base.sort!(ctx::Context2D3V, s::Integer, Δx::Real, Δy::Real)
    p = zeros(Int64, ctx.cfg.species[s].P)
    rc = GC.@preserve p ccall((:pic2d3v_sort, libpicgolf), Cint, (Ptr{Cvoid}, Int32, Float64, Float64, Ptr{Int64}),
                              ctx.ptr, s - 1, Δx, Δy, p)
    check(ctx, rc, "pic2d3v_sort")
    p .+ 1
end

"""
    run!(ctx::Context2D3V, nsteps) -> rows

`nsteps` of `loop!(plasma, field, ...); diagnose!(d, field, plasma, t, ...)`;
one row per `t % ntskip == 0` of this call: kineticenergy, fieldenergy,
particlemomentum (3), characteristicmomentum (3); the Lorenz gauge adds
fieldmomentum (3) before characteristicmomentum.
"""
function run!(ctx::Context2D3V, nsteps::Integer)
    nt = ctx.cfg.ntskip
    rows = cld(ctx.t + nsteps, nt) - cld(ctx.t, nt)
    R = zeros(ctx.cfg.field in (FIELDS2D3V[:lorenz_staggered], FIELDS2D3V[:lorenz_semi_implicit]) ? 11 : 8, rows)
    rc = GC.@preserve R ccall((:pic2d3v_run, libpicgolf), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}),
                              ctx.ptr, nsteps, rows > 0 ? pointer(R) : C_NULL)
    ctx.t += nsteps
    check(ctx, rc, "pic2d3v_run")
    permutedims(R)
end

"`fields(ctx) -> (Exs, Eys, ϕs)`: the diagnostics' binned fields (NX÷ngskip x NY÷ngskip x bins)."
function fields(ctx::Context2D3V)
    c = ctx.cfg
    nb = cld(ctx.t, c.ntskip)
    A = [zeros(c.NX ÷ c.ngskip, c.NY ÷ c.ngskip, nb) for _ in 1:3]
    if nb > 0
        rc = GC.@preserve A ccall((:pic2d3v_fields, libpicgolf), Cint,
                                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64),
                                  ctx.ptr, A[1], A[2], A[3], nb)
        check(ctx, rc, "pic2d3v_fields")
    end
    (A[1], A[2], A[3])
end

"`field(ctx) -> (Ex, Ey)`: the gathered field `Exy` (interior NX x NY)."
function field(ctx::Context2D3V)
    Ex, Ey = zeros(ctx.cfg.NX, ctx.cfg.NY), zeros(ctx.cfg.NX, ctx.cfg.NY)
    rc = GC.@preserve Ex Ey ccall((:pic2d3v_field, libpicgolf), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                                  ctx.ptr, Ex, Ey)
    check(ctx, rc, "pic2d3v_field")
    (Ex, Ey)
end

"`deposit!(ρ, ctx, shape, x, y, w)`: ρ = Σ (wx*wy)*w over the particles, folded periodically."
function deposit!(ρ::Matrix{Float64}, ctx::Context2D3V, shape::Symbol, x::Vector{Float64}, y::Vector{Float64},
                  w::Real)
    rc = GC.@preserve ρ x y ccall((:pic2d3v_deposit, libpicgolf), Cint,
                                  (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Ptr{Float64}),
                                  ctx.ptr, SHAPES2D3V[shape], x, y, length(x), w, ρ)
    check(ctx, rc, "pic2d3v_deposit")
    ρ
end

"`gather(ctx, shape, Ex, Ey, x, y) -> (ex, ey)`: `field(shape, x[j], y[j])` for every j."
function gather(ctx::Context2D3V, shape::Symbol, Ex::Matrix{Float64}, Ey::Matrix{Float64}, x::Vector{Float64},
                y::Vector{Float64})
    ex, ey = zeros(length(x)), zeros(length(x))
    rc = GC.@preserve Ex Ey x y ex ey ccall((:pic2d3v_gather, libpicgolf), Cint,
                                            (Ptr{Cvoid}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                                             Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}),
                                            ctx.ptr, SHAPES2D3V[shape], Ex, Ey, x, y, length(x), ex, ey)
    check(ctx, rc, "pic2d3v_gather")
    (ex, ey)
end

"`boris!(ctx, vx, vy, vz, Ex, Ey; q_m)`: the context's Boris (B0, dt, variant) in place."
function boris!(ctx::Context2D3V, vx::Vector{Float64}, vy::Vector{Float64}, vz::Vector{Float64},
                Ex::Vector{Float64}, Ey::Vector{Float64}; q_m::Real)
    rc = GC.@preserve vx vy vz Ex Ey ccall((:pic2d3v_boris, libpicgolf), Cint,
                                           (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                                            Ptr{Float64}, Int64, Float64),
                                           ctx.ptr, vx, vy, vz, Ex, Ey, length(vx), q_m)
    check(ctx, rc, "pic2d3v_boris")
    (vx, vy, vz)
end

"`warmup!(ctx::Context2D3V)`: `warmup!(field, plasma, to)` (nothing for an electrostatic field)."
function warmup!(ctx::Context2D3V)
    check(ctx, ccall((:pic2d3v_warmup, libpicgolf), Cint, (Ptr{Cvoid},), ctx.ptr), "pic2d3v_warmup")
    ctx
end

"`loop!(ctx::Context2D3V, nsteps)`: `nsteps` of the Lorenz-gauge `loop!` without `diagnose!`."
function loop!(ctx::Context2D3V, nsteps::Integer)
    rc = ccall((:pic2d3v_loop, libpicgolf), Cint, (Ptr{Cvoid}, Int64), ctx.ptr, nsteps)
    check(ctx, rc, "pic2d3v_loop")
    ctx.t += nsteps
    ctx
end

"`lg_field(ctx, which) -> Matrix{ComplexF64}`: the Lorenz-gauge field array `which` (a symbol of `LG_STATE`)."
function lg_field(ctx::Context2D3V, which::Symbol)
    z = zeros(ComplexF64, ctx.cfg.NX, ctx.cfg.NY)
    k = Int32(findfirst(==(which), LG_STATE) - 1)
    rc = GC.@preserve z ccall((:pic2d3v_lg_get, libpicgolf), Cint, (Ptr{Cvoid}, Int32, Ptr{ComplexF64}),
                              ctx.ptr, k, z)
    check(ctx, rc, "pic2d3v_lg_get")
    z
end

"`lg_field!(ctx, which, z)`: sets the Lorenz-gauge field array `which`."
function lg_field!(ctx::Context2D3V, which::Symbol, z::Matrix{ComplexF64})
    k = Int32(findfirst(==(which), LG_STATE) - 1)
    rc = GC.@preserve z ccall((:pic2d3v_lg_set, libpicgolf), Cint, (Ptr{Cvoid}, Int32, Ptr{ComplexF64}),
                              ctx.ptr, k, z)
    check(ctx, rc, "pic2d3v_lg_set")
    ctx
end

"""`lg_grid(ctx) -> Array{Float64,3}`: a `LorenzGaugeSemiImplicitField`'s deposit grid `ρJs⁺`
(4 x (NX+2buffer) x (NY+2buffer), axes 1:4, -(buffer-1):NX+buffer, ... as the reference's)."""
function lg_grid(ctx::Context2D3V)
    c = ctx.cfg
    A = zeros(4, c.NX + 2c.buffer, c.NY + 2c.buffer)
    rc = GC.@preserve A ccall((:pic2d3v_lg_grid, libpicgolf), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.ptr, A)
    check(ctx, rc, "pic2d3v_lg_grid")
    A
end

"`lg_grid!(ctx, A)`: sets the deposit grid `ρJs⁺` (the next step's first convergence test compares with it)."
function lg_grid!(ctx::Context2D3V, A::Array{Float64,3})
    rc = GC.@preserve A ccall((:pic2d3v_lg_set_grid, libpicgolf), Cint, (Ptr{Cvoid}, Ptr{Float64}), ctx.ptr, A)
    check(ctx, rc, "pic2d3v_lg_set_grid")
    ctx
end

"`lg_iterations(ctx, n) -> Vector{Int64}`: the semi-implicit iteration counts of the last `n` steps."
function lg_iterations(ctx::Context2D3V, n::Integer)
    v = zeros(Int64, max(n, 1))
    rc = GC.@preserve v ccall((:pic2d3v_lg_iterations, libpicgolf), Cint, (Ptr{Cvoid}, Ptr{Int64}, Int64),
                              ctx.ptr, v, n)
    check(ctx, rc, "pic2d3v_lg_iterations")
    v[1:n]
end

"`lg_bins(ctx, which) -> Array{Float64,3}`: `LorenzGaugeDiagnostics`' binned field `which` (e.g. `:Ezs`)."
function lg_bins(ctx::Context2D3V, which::Symbol)
    c = ctx.cfg
    nb = cld(ctx.t, c.ntskip)
    A = zeros(c.NX ÷ c.ngskip, c.NY ÷ c.ngskip, nb)
    k = Int32(findfirst(==(which), LG_BINS) - 1)
    if nb > 0
        rc = GC.@preserve A ccall((:pic2d3v_lg_bins, libpicgolf), Cint, (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64),
                                  ctx.ptr, k, A, nb)
        check(ctx, rc, "pic2d3v_lg_bins")
    end
    A
end

"`fieldsolve!(Ex, Ey, ctx, ρ; ϕ=nothing)`: the electrostatic solve of `loop!` in the context's box."
function fieldsolve!(Ex::Matrix{Float64}, Ey::Matrix{Float64}, ctx::Context2D3V, ρ::Matrix{Float64};
                     ϕ::Union{Matrix{Float64},Nothing}=nothing)
    rc = GC.@preserve Ex Ey ρ ϕ ccall((:pic2d3v_fieldsolve, libpicgolf), Cint,
                                      (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                      ctx.ptr, ρ, Ex, Ey, ϕ === nothing ? C_NULL : ϕ)
    check(ctx, rc, "pic2d3v_fieldsolve")
    (Ex, Ey)
end

# ---------------------------------------------------------------------------
# 1D electromagnetic fixed point (src/GaussianEMFixedPoint.jl, scheme
# :fixed_point) and explicit step (src/GaussianEMExplicit.jl, scheme
# :explicit), C ABI in include/picem1d.h.
#
# | script expression                                          | PICGolf                              |
# |------------------------------------------------------------|--------------------------------------|
# | `d(y, u)` (line 11)                                          | `deposit(ctx, y, w; u)`              |
# | one sweep's `Ex .= ...` .. `bz=(Bz.+Cz)/2` (lines 20-27)     | `fields!(ctx, ρ, Jy, Jz, F, C, E, B)` |
# | `v[j,:] .= u .+ (a .+ cross(u .+ a*dt/2, b))*dt` (28-29)     | `push(ctx, c, e3, b2, u)`            |
# | the `@gif for t ∈ 1:T` loop and `fields` (lines 13-50)       | `upload!`, `run!`, `download`        |
# | explicit: Gauss, Faraday, Ampere (lines 13, 15-16, 23-24)    | `explicit_fields!(ctx, ρ, Jy, Jz, E, B)` |
# | explicit: `v[j,:].+=a*dt/2` .. (lines 17-22)                 | `explicit_push(ctx, x, E, b2, u)`    |
# | explicit: the `for t ∈ 1:T` loop and `fields` (lines 12-30)  | `upload!`, `run!`, `download`        |
# ---------------------------------------------------------------------------

const EM_SCHEMES = Dict(:fixed_point => Int32(0), :explicit => Int32(1))

"Mirror of `struct picem1d_config` in include/picem1d.h."
struct PicEM1DConfig
    N::Int32
    max_sweeps::Int32
    P::Int64
    eps::Float64
    dt::Float64
    w::Float64
    rtol::Float64
    sort::Int32
    device::Int32
    scheme::Int32
end

"""
    PicEM1DConfig(; N, P, eps, dt, w, rtol=1e-8, max_sweeps=3, sort=true, device=0, scheme=:fixed_point)

The script's N, P, ϵ, dt = ϵ^2/N/2, w = W/P*N, η, the three sweeps of
`for _∈0:2`, and the `sortperm(x)` after every step; `scheme=:explicit` runs
src/GaussianEMExplicit.jl instead (η and the sweeps are then unused).
"""
PicEM1DConfig(; N, P, eps, dt, w, rtol=1e-8, max_sweeps=3, sort=true, device=0, scheme=:fixed_point) =
    PicEM1DConfig(Int32(N), Int32(max_sweeps), Int64(P), Float64(eps), Float64(dt), Float64(w), Float64(rtol),
                  Int32(sort), Int32(device), EM_SCHEMES[scheme])

mutable struct ContextEM1D
    ptr::Ptr{Cvoid}
    cfg::PicEM1DConfig
end

function check(ctx::ContextEM1D, rc::Cint, what::AbstractString)
    rc == 0 || error("$what: ", unsafe_string(ccall((:picem1d_last_error, libpicgolf), Cstring, (Ptr{Cvoid},), ctx.ptr)))
    nothing
end

function ContextEM1D(cfg::PicEM1DConfig)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:picem1d_create, libpicgolf), Cint, (Ref{Ptr{Cvoid}}, Ref{PicEM1DConfig}), h, cfg)
    rc == 0 || error("picem1d_create: ", unsafe_string(ccall((:picem1d_last_error, libpicgolf), Cstring,
                                                             (Ptr{Cvoid},), C_NULL)))
    ctx = ContextEM1D(h[], cfg)
    finalizer(destroy!, ctx)
    ctx
end

function destroy!(ctx::ContextEM1D)
    if ctx.ptr != C_NULL
        ccall((:picem1d_destroy, libpicgolf), Cint, (Ptr{Cvoid},), ctx.ptr)
        ctx.ptr = C_NULL
    end
    nothing
end

"""
    upload!(ctx::ContextEM1D, x, v; E=nothing, B=nothing)

The script's `x` and `v` (P x 3); `E` is N x 3 (Ex, Ey, Ez; zeros), `B` is
N x 2 (By, Bz; the script's zeros and ones).  Restarts the script's Fz.
"""
function upload!(ctx::ContextEM1D, x::Vector{Float64}, v::Matrix{Float64};
                 E::Union{Matrix{Float64},Nothing}=nothing, B::Union{Matrix{Float64},Nothing}=nothing)
    rc = GC.@preserve x v E B ccall((:picem1d_upload, libpicgolf), Cint,
                                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                    ctx.ptr, x, v, E === nothing ? C_NULL : E, B === nothing ? C_NULL : B)
    check(ctx, rc, "picem1d_upload")
    ctx
end

"`download(ctx::ContextEM1D) -> (x, v, E, B)` (v: P x 3, E: N x 3, B: N x 2)."
function download(ctx::ContextEM1D)
    P, N = ctx.cfg.P, ctx.cfg.N
    x, v, E, B = zeros(P), zeros(P, 3), zeros(N, 3), zeros(N, 2)
    rc = GC.@preserve x v E B ccall((:picem1d_download, libpicgolf), Cint,
                                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                                    ctx.ptr, x, v, E, B)
    check(ctx, rc, "picem1d_download")
    (x, v, E, B)
end

"""
    run!(ctx::ContextEM1D, T) -> (fields, sweeps)

`T` steps of the loop; `fields` is the script's `fields` (N x T x 12) for
these steps.
"""
function run!(ctx::ContextEM1D, T::Integer)
    f = zeros(ctx.cfg.N, T, 12)
    sw = zeros(Int32, T)
    rc = GC.@preserve f sw ccall((:picem1d_run, libpicgolf), Cint, (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Int32}),
                                 ctx.ptr, T, f, sw)
    check(ctx, rc, "picem1d_run")
    (f, sw)
end

"`deposit(ctx::ContextEM1D, y, w; u=nothing) -> r`: `d(y, u)`."
function deposit(ctx::ContextEM1D, y::Vector{Float64}, w::Real; u::Union{Vector{Float64},Nothing}=nothing)
    r = zeros(ctx.cfg.N)
    rc = GC.@preserve y u r ccall((:picem1d_deposit, libpicgolf), Cint,
                                  (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Ptr{Float64}),
                                  ctx.ptr, y, u === nothing ? C_NULL : u, length(y), w, r)
    check(ctx, rc, "picem1d_deposit")
    r
end

"""
    fields!(ctx::ContextEM1D, ρ, Jy, Jz, Fx, Fy, Fz, Cy, Cz, E, B) -> (e3, b2)

One sweep's updates: `E` (N x 3) and `B` (N x 2, the iterate's By, Bz) in
place; returns `ex, ey, ez` (N x 3) and `by, bz` (N x 2).
"""
function fields!(ctx::ContextEM1D, ρ, Jy, Jz, Fx, Fy, Fz, Cy, Cz, E::Matrix{Float64}, B::Matrix{Float64})
    N = ctx.cfg.N
    e3, b2 = zeros(N, 3), zeros(N, 2)
    rc = GC.@preserve ρ Jy Jz Fx Fy Fz Cy Cz E B e3 b2 ccall((:picem1d_fields, libpicgolf), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        ctx.ptr, ρ, Jy, Jz, Fx, Fy, Fz, Cy, Cz, E, B, e3, b2)
    check(ctx, rc, "picem1d_fields")
    (e3, b2)
end

"`push(ctx::ContextEM1D, c, e3, b2, u) -> v` (u, v: n x 3): the velocity update at midpoints c."
function push(ctx::ContextEM1D, c::Vector{Float64}, e3::Matrix{Float64}, b2::Matrix{Float64}, u::Matrix{Float64})
    v = zeros(size(u))
    rc = GC.@preserve c e3 b2 u v ccall((:picem1d_push, libpicgolf), Cint,
                                        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64,
                                         Ptr{Float64}), ctx.ptr, c, e3, b2, u, length(c), v)
    check(ctx, rc, "picem1d_push")
    v
end

"""
    explicit_fields!(ctx::ContextEM1D, ρ, Jy, Jz, E, B) -> b2

The explicit step's field updates: `Ex` from `ρ` (line 13), `By, Bz` from the
given `Ey, Ez` (line 15), then `Ey, Ez` from the new `By, Bz` and `Jy, Jz`
(lines 23-24), with `E` (N x 3) and `B` (N x 2) in place; returns `by, bz`
(N x 2) of line 16.
"""
function explicit_fields!(ctx::ContextEM1D, ρ, Jy, Jz, E::Matrix{Float64}, B::Matrix{Float64})
    b2 = zeros(ctx.cfg.N, 2)
    rc = GC.@preserve ρ Jy Jz E B b2 ccall((:picem1d_explicit_fields, libpicgolf), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        ctx.ptr, ρ, Jy, Jz, E, B, b2)
    check(ctx, rc, "picem1d_explicit_fields")
    b2
end

"`explicit_push(ctx::ContextEM1D, x, E, b2, u) -> v` (u, v: n x 3): the split kick of lines 17-22 at x."
function explicit_push(ctx::ContextEM1D, x::Vector{Float64}, E::Matrix{Float64}, b2::Matrix{Float64},
                       u::Matrix{Float64})
    v = zeros(size(u))
    rc = GC.@preserve x E b2 u v ccall((:picem1d_explicit_push, libpicgolf), Cint,
                                       (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64,
                                        Ptr{Float64}), ctx.ptr, x, E, b2, u, length(x), v)
    check(ctx, rc, "picem1d_explicit_push")
    v
end

end # module
