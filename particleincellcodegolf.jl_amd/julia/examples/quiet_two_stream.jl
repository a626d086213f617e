# Quiet-start two-stream run (the physics of src/GaussianFixedPointQuiet.jl)
# through PICGolf: the whole fixed-point loop runs on the GPU, the diagnostics
# come back as the script's D matrix.
include(joinpath(@__DIR__, "..", "PICGolf.jl"))
using .PICGolf

N = 64; P = 32N; dt = 1 / 6N; T = 2^13; W = 32π^2 / 3
cfg = PicConfig(N=N, P=P, W=W, dt=dt, shape=:gauss7, scheme=:midpoint, rtol=4eps(), max_sweeps=10)
ctx = Context(cfg)
x, v = quiet_start(P)
upload!(ctx, x, v)                 # E starts at zeros(N) as in the script
D, sweeps = run!(ctx, T)           # D[t, 1:4, 1] == the script's D[t, 1:4]
γ(k) = imag(sqrt(Complex(k^2 + 1 - sqrt(4k^2 + 1)))) * sqrt(W / 2) / log(10)
println("field energy: ", D[1, 1, 1], " -> ", D[end, 1, 1], "; predicted slope ", 2γ(2π / sqrt(W / 2)))

# operator-level use, mirroring one Picard sweep of the script
X, V = copy(x), copy(v); E = zeros(N); r = zeros(N)
xm = midpoint_position!(similar(x), ctx, X, V, v, dt)
deposit!(r, ctx, xm, X)
fieldsolve!(E, ctx, r)
push!(ctx, X, V, xm, v, E, dt)
