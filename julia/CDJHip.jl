# CDJHip.jl - ccall binding of libcdjhip.so (include/cdj.h) for CutAndDisplaceJulia.
#
# Drop-in replacements for TD (src/TD.jl:77, :135, :219), CalculateInfluenceMatrices3D
# (src/CalculateInfluenceMatrices3D.jl:2), SlipCalculator3D (src/SlipCalculator3D.jl:1),
# GetCrackTipElements3D (src/GetCrackTipElements3D.jl:1) and StressIntensity3D (src/StressIntensity3D.jl:1) with
# the same signatures, and PolygonRemeshingCGAL for BuildCGAL's (the growth loops' remeshing step).  Include this file after CutAndDisplaceJulia and call CDJHip.TD(...),
# CDJHip.CalculateInfluenceMatrices3D(...) and CDJHip.SlipCalculator3D(...) where the package's functions were
# called.  The precision mode lives in module-level settings (CDJHip.PRECISION[], CDJHip.ASSEMBLY[]), so the
# public signatures stay unchanged.
module CDJHip

using ..CutAndDisplaceJulia: StrainInf, DispInf, Strains, Disps, TractionInf, SetupTractionVector, KnockOutFixedRows,
    OFFReader, OFFExport,
    TriangleEdges, MixedBoundaryConditionsFriction, MixedBoundaryConditionsFluidVolume

const LIB = get(ENV, "CDJ_LIB", joinpath(@__DIR__, "..", "cutanddisplacejulia_amd", "libcdjhip.so"))
const CTX = Ref{Ptr{Cvoid}}(C_NULL)
const PRECISION = Ref{Int32}(0)        # 0 = FP64 (reference arithmetic), 1 = FP32 LU + FP64 GMRES-IR
const ASSEMBLY = Ref{Int32}(0)         # 0 = FP64 assembly, 1 = FP32 far field / FP64 near field
const NEAR_FACTOR = Ref{Float64}(2.0)  # FP64 radius in longest triangle edges about each source centroid
const LOCK = ReentrantLock()

struct Opts
    precision::Int32
    tile::Int32
    tol::Float64
    max_iter::Int32
    assembly::Int32
    near_factor::Float64
end
Opts() = Opts(PRECISION[], 0, 1e-10, 0, ASSEMBLY[], NEAR_FACTOR[])

mutable struct Report
    path::Int32
    iterations::Int32
    residual::Float64
    assembly_ms::Float64
    solve_ms::Float64
    near_entries::Int64
end
Report() = Report(0, 0, 0.0, 0.0, 0.0, 0)

function ctx()
    if CTX[] == C_NULL
        h = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:cdj_create, LIB), Cint, (Cint, Ref{Ptr{Cvoid}}), 0, h)
        rc == 0 || error("cdj_create failed: " * unsafe_string(ccall((:cdj_last_error, LIB), Cstring, (Ptr{Cvoid},), C_NULL)))
        CTX[] = h[]
    end
    return CTX[]
end

check(rc) = rc == 0 || error(unsafe_string(ccall((:cdj_last_error, LIB), Cstring, (Ptr{Cvoid},), ctx())))

_f64(a) = convert(Array{Float64}, a)

function TD(X, Y, Z, P1List, P2List, P3List, Dss, Dds, Dn, ν, G, DispFlag, StrainFlag, HSflag)
    n = length(P1List) == 3 ? 1 : size(P1List, 1)
    m = length(X)
    S = StrainFlag == 1 ? StrainInf((zeros(m, n) for _ in 1:18)...) : StrainInf(([] for _ in 1:18)...)
    D = DispFlag == 1 ? DispInf((zeros(m, n) for _ in 1:9)...) : DispInf(([] for _ in 1:9)...)
    return TD(X, Y, Z, P1List, P2List, P3List, Dss, Dds, Dn, ν, G, DispFlag, StrainFlag, HSflag, S, D)
end

function TD(X, Y, Z, P1List, P2List, P3List, Dss, Dds, Dn, ν, G, DispFlag, StrainFlag, HSflag,
            S::StrainInf, D::DispInf)
    P1 = _f64(reshape(P1List, :, 3)); P2 = _f64(reshape(P2List, :, 3)); P3 = _f64(reshape(P3List, :, 3))
    x = _f64(vec(X)); y = _f64(vec(Y)); z = _f64(vec(Z))
    sp = StrainFlag == 1 ? [pointer(getfield(S, k)) for k in fieldnames(StrainInf)] : Ptr{Float64}[]
    dp = DispFlag == 1 ? [pointer(getfield(D, k)) for k in fieldnames(DispInf)] : Ptr{Float64}[]
    lock(LOCK) do
        GC.@preserve S D P1 P2 P3 x y z sp dp begin
            check(ccall((:cdj_td_influence, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Int32, Int32, Int32,
                 Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}),
                ctx(), x, y, z, length(x), P1, P2, P3, size(P1, 1), _f64(vec(Dss)), _f64(vec(Dds)), _f64(vec(Dn)),
                ν, G, DispFlag, StrainFlag, HSflag,
                isempty(sp) ? C_NULL : sp, isempty(dp) ? C_NULL : dp))
        end
    end
    return (S, D)
end

function TD(X, Y, Z, P1List, P2List, P3List, Dss, Dds, Dn, ν, G, DispFlag, StrainFlag, HSflag,
            ::Strains, ::Disps)
    P1 = _f64(reshape(P1List, :, 3)); P2 = _f64(reshape(P2List, :, 3)); P3 = _f64(reshape(P3List, :, 3))
    x = _f64(vec(X)); y = _f64(vec(Y)); z = _f64(vec(Z)); m = length(x)
    s6 = [zeros(m) for _ in 1:6]; d3 = [zeros(m) for _ in 1:3]
    sp = [pointer(a) for a in s6]; dp = [pointer(a) for a in d3]
    lock(LOCK) do
        GC.@preserve s6 d3 sp dp P1 P2 P3 x y z begin
            check(ccall((:cdj_td_field, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
                 Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Int32, Int32, Int32,
                 Ptr{Ptr{Float64}}, Ptr{Ptr{Float64}}),
                ctx(), x, y, z, m, P1, P2, P3, size(P1, 1), _f64(vec(Dss)), _f64(vec(Dds)), _f64(vec(Dn)),
                ν, G, DispFlag, StrainFlag, HSflag, StrainFlag == 1 ? sp : C_NULL, DispFlag == 1 ? dp : C_NULL))
        end
    end
    rs(a) = size(X, 2) != 1 ? reshape(a, size(X)) : a
    S = StrainFlag == 1 ? Strains(rs.(s6)...) : Strains([], [], [], [], [], [])
    D = DispFlag == 1 ? Disps(rs.(d3)...) : Disps([], [], [])
    return (S, D)
end

# FixedEls as the library's Int32 flags (1 = fixed), C_NULL when no element is fixed
_fixed(FixedEls) = any(FixedEls .== 1) ? Int32.(vec(FixedEls) .== 1) : C_NULL

function CalculateInfluenceMatrices3D(FaceNormalVector, MidPoint, P1, P2, P3, ν, G, λ, FixedEls, HSFlag, n)
    fx = _fixed(FixedEls)
    nf = fx === C_NULL ? n : n - sum(fx)
    T = [zeros(n, nf) for _ in 1:9]       # N x nf with fixed elements (CalculateInfluenceMatrices3D.jl:10-33)
    tp = [pointer(a) for a in T]
    opts = Ref(Opts())
    nrm = _f64(FaceNormalVector); mid = _f64(MidPoint)
    lock(LOCK) do
        GC.@preserve T tp nrm mid fx begin
            check(ccall((:cdj_traction_influence, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64,
                 Float64, Float64, Int32, Ptr{Int32}, Ptr{Ptr{Float64}}, Ref{Opts}),
                ctx(), nrm, mid, _f64(P1), _f64(P2), _f64(P3), n, ν, G, λ, HSFlag, fx, tp, opts))
        end
    end
    return TractionInf(T...)
end

function SlipCalculator3D(P1, P2, P3, ν, G, λ, MidPoint, FaceNormalVector, HSFlag, BoundaryConditions, FixedEls)
    n = size(FaceNormalVector, 1)
    (Tn, Tds, Tss) = SetupTractionVector(BoundaryConditions, FaceNormalVector, n, G, λ)
    (Tn, Tds, Tss) = KnockOutFixedRows(FixedEls, Tn, Tds, Tss)
    b = [Tn; Tss; Tds]
    fx = _fixed(FixedEls)
    nf = fx === C_NULL ? n : n - sum(fx)  # with fixed elements A is 3n x 3nf and x its least-squares solution
    x = zeros(3nf); A = zeros(3n, 3nf)
    opts = Ref(Opts()); rep = Ref(Report())
    lock(LOCK) do
        GC.@preserve fx begin
            check(ccall((:cdj_slip_calculator, LIB), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Float64, Float64, Ptr{Float64},
                 Ptr{Float64}, Int32, Ptr{Float64}, Int64, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Ref{Opts},
                 Ref{Report}),
                ctx(), _f64(P1), _f64(P2), _f64(P3), n, ν, G, λ, _f64(MidPoint), _f64(FaceNormalVector), HSFlag,
                _f64(b), 1, fx, x, A, opts, rep))
        end
    end
    return x[1:nf], x[nf+1:2nf], x[2nf+1:3nf], A, b
end

function SlipCalculator3D(P1, P2, P3, ν, G, λ, MidPoint, FaceNormalVector, HSFlag,
                          BoundaryConditions::MixedBoundaryConditionsFriction, FixedEls)
    n = size(FaceNormalVector, 1)
    any(FixedEls .== 1) && error("fixed elements: use the reference implementation")
    (Tn, Tds, Tss) = SetupTractionVector(BoundaryConditions.MixedBoundaryConditions, FaceNormalVector, n, G, λ)
    b = [Tn; Tss; Tds]
    µ = zeros(n) .+ BoundaryConditions.FrictionParameters.µ
    Sf = zeros(n) .+ BoundaryConditions.FrictionParameters.Sf
    Dn = zeros(n); Dss = zeros(n); Dds = zeros(n)
    rep = Ref(Report())
    lock(LOCK) do
        check(ccall((:cdj_slip_friction, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Float64, Float64, Ptr{Float64},
             Ptr{Float64}, Int32, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Ref{Report}),
            ctx(), _f64(P1), _f64(P2), _f64(P3), n, ν, G, λ, _f64(MidPoint), _f64(FaceNormalVector), HSFlag,
            _f64(b), µ, Sf, Dn, Dss, Dds, rep))
    end
    return Dn, Dss, Dds
end

function SlipCalculator3D(P1, P2, P3, ν, G, λ, MidPoint, FaceNormalVector, HSFlag,
                          BoundaryConditions::MixedBoundaryConditionsFluidVolume, FractureFlag)
    n = size(FaceNormalVector, 1)
    (Tn, Tds, Tss) = SetupTractionVector(BoundaryConditions.MixedBoundaryConditions, FaceNormalVector, n, G, λ)
    b = [Tn; Tss; Tds]
    flag = Int32.(vec(FractureFlag))
    V = Float64(BoundaryConditions.Volumes[1])
    Dn = zeros(n); Dss = zeros(n); Dds = zeros(n); pressure = Ref(0.0)
    rep = Ref(Report())
    lock(LOCK) do
        check(ccall((:cdj_slip_fluid_volume, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Float64, Float64, Ptr{Float64},
             Ptr{Float64}, Int32, Ptr{Float64}, Ptr{Int32}, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
             Ref{Float64}, Ref{Report}),
            ctx(), _f64(P1), _f64(P2), _f64(P3), n, ν, G, λ, _f64(MidPoint), _f64(FaceNormalVector), HSFlag,
            _f64(b), flag, V, Dn, Dss, Dds, pressure, rep))
    end
    return Dn, Dss, Dds
end

# Edge blocks P1P2, P1P3, P2P3; geometry columns FeLe, FeMd (3), FeEv (3), FeM2Ev (3), FeM2ELe, IntAng, Area
function GetCrackTipElements3D(MidPoint, P1, P2, P3, FaceNormalVector)
    n = size(P1, 1)
    flags = zeros(UInt8, 3n); geom = zeros(3n * 13)
    lock(LOCK) do
        check(ccall((:cdj_crack_tip_elements, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Ptr{UInt8},
             Ptr{Float64}),
            ctx(), _f64(MidPoint), _f64(P1), _f64(P2), _f64(P3), _f64(FaceNormalVector), n, flags, geom))
    end
    blocks = map(0:2) do s
        g = reshape(geom[s*n*13+1:(s+1)*n*13], n, 13)
        TriangleEdges(g[:, 1], g[:, 2:4], g[:, 5:7], g[:, 8:10], flags[s*n+1:(s+1)*n] .== 1, g[:, 11], g[:, 12],
                      0.0, 0.0, 0.0, 0.0, g[:, 13])
    end
    return blocks[1], blocks[2], blocks[3]
end

function StressIntensity3D(Dn, Dss, Dds, G, ν, FaceNormalVector, FeP1P2S, FeP1P3S, FeP2P3S)
    n = size(FaceNormalVector, 1)
    E = (FeP1P2S, FeP1P3S, FeP2P3S)
    flags = vcat([UInt8.(e.FreeFlg) for e in E]...)
    geom = vcat([vec(hcat(e.FeLe, e.FeMd, e.FeEv, e.FeM2Ev, e.FeM2ELe, e.IntAng, e.Area)) for e in E]...)
    K = zeros(3n * 4)
    lock(LOCK) do
        check(ccall((:cdj_stress_intensity, LIB), Cint,
            (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int64, Float64, Float64, Ptr{Float64}, Ptr{UInt8},
             Ptr{Float64}, Ptr{Float64}),
            ctx(), _f64(vec(Dn)), _f64(vec(Dss)), _f64(vec(Dds)), n, G, ν, _f64(FaceNormalVector), flags, geom, K))
    end
    for (s, e) in enumerate(E)
        k = reshape(K[(s-1)*n*4+1:s*n*4], n, 4)
        e.K1 = k[:, 1]; e.K2 = k[:, 2]; e.K3 = k[:, 3]; e.StrainEnergy = k[:, 4]
    end
    return FeP1P2S, FeP1P3S, FeP2P3S
end

# Isotropic remeshing in place of BuildCGAL.PolygonRemeshingCGAL (file in, file out, as the reference's growth loops
# call it between OFFExport and OFFReader; test/test_PropagationTest.jl:151-561).  Host code, no GPU.
function PolygonRemeshingCGAL(OutputDirectory, target_edge_length; iterations=3)
    (Points, Triangles) = OFFReader(OutputDirectory)
    pos = Dict(Int(Points[i, 1]) => i for i in 1:size(Points, 1))
    tris = Int64[pos[Int(t)] - 1 for t in Triangles]          # 0-based rows of Points
    pts = convert(Matrix{Float64}, Points[:, 2:4])
    h = Ref{Ptr{Cvoid}}(C_NULL); np = Ref{Int64}(0); nt = Ref{Int64}(0)
    check(ccall((:cdj_remesh_isotropic, LIB), Cint,
        (Ptr{Float64}, Int64, Ptr{Int64}, Int64, Float64, Int32, Ref{Ptr{Cvoid}}, Ref{Int64}, Ref{Int64}),
        pts, size(pts, 1), tris, size(tris, 1), target_edge_length, iterations, h, np, nt))
    P = zeros(np[], 3); T = zeros(Int64, nt[], 3)
    check(ccall((:cdj_remesh_take, LIB), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Int64}), h[], P, T))
    Pts = hcat(collect(1.0:np[]), P)
    base = splitext(OutputDirectory)[1] * "-Remeshed"   # absolute: the package's OFFExport would prefix pwd()
    OFFExport(Pts, T .+ 1, nt[], np[], base)
    return base * ".off"
end

end # module
