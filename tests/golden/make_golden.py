"""Writes tests/golden/ref_sampling.npz and ref_flake.npz: input/output
vectors of the reference's own sampling warps, Fresnel terms, basis
construction, Matrix3 operations and SGGX micro-flake phase sampling.

Run in a checkout that has the reference sources, after `make -C oracle ref`
has compiled them into oracle/_ref/libtempest_ref.so (see oracle/ref_shim.cc
for the functions and their file:line):

    python tests/golden/make_golden.py

The fixture holds data only: seeded inputs and the reference's outputs.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libtempest_ref.so"))
    F = ctypes.c_float
    P = ctypes.POINTER(F)
    for name in ("ref_fresnel_conductor", "ref_fresnel_schlick", "ref_fresnel_dielectric", "ref_cos_transmittance",
                 "ref_sphere_area_pdf", "ref_projected_triangle_area"):
        getattr(lib, name).restype = F
    rng = np.random.Generator(np.random.PCG64(2024))
    n = 512

    def fp(a):
        return a.ctypes.data_as(P)

    out = {}
    uv = rng.uniform(0.0, 1.0, size=(n, 2)).astype(np.float32)
    uv[:8] = [[0, 0], [1, 0], [0, 1], [1, 1], [0.5, 0.5], [0.25, 0.75], [1e-6, 0.999999], [0.999999, 1e-6]]
    out["warp_uv"] = uv
    for name in ("ref_cosine_sample_hemisphere", "ref_uniform_sample_hemisphere", "ref_uniform_sample_sphere"):
        res = np.zeros((n, 3), dtype=np.float32)
        for i in range(n):
            getattr(lib, name)(F(uv[i, 0]), F(uv[i, 1]), fp(res[i]))
        out[name[4:]] = res

    # sphere-area light sampling and PDF
    pos = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
    cen = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
    rad = rng.uniform(0.05, 0.6, size=n).astype(np.float32)
    sdir = np.zeros((n, 3), dtype=np.float32)
    sdist = np.zeros(n, dtype=np.float32)
    spdf = np.zeros(n, dtype=np.float32)
    qpdf = np.zeros(n, dtype=np.float32)
    for i in range(n):
        d, pp = F(), F()
        lib.ref_sample_sphere_area(fp(pos[i]), fp(cen[i]), F(rad[i]), F(uv[i, 0]), F(uv[i, 1]), fp(sdir[i]),
                                   ctypes.byref(d), ctypes.byref(pp))
        sdist[i], spdf[i] = d.value, pp.value
        qpdf[i] = lib.ref_sphere_area_pdf(fp(pos[i]), fp(sdir[i]), fp(cen[i]), F(rad[i]))
    out.update(sphere_pos=pos, sphere_center=cen, sphere_radius=rad, sphere_dir=sdir, sphere_dist=sdist,
               sphere_pdf=spdf, sphere_pdf_eval=qpdf)

    # triangle-area light sampling and projected area
    tri = rng.uniform(-1.0, 1.0, size=(n, 3, 3)).astype(np.float32)
    tpos = rng.uniform(-1.0, 1.0, size=(n, 3)).astype(np.float32)
    tdir = np.zeros((n, 3), dtype=np.float32)
    tdist = np.zeros(n, dtype=np.float32)
    tpdf = np.zeros(n, dtype=np.float32)
    parea = np.zeros(n, dtype=np.float32)
    for i in range(n):
        d, pp = F(), F()
        lib.ref_sample_triangle_area(fp(tpos[i]), fp(tri[i, 0]), fp(tri[i, 1]), fp(tri[i, 2]), F(uv[i, 0]),
                                     F(uv[i, 1]), fp(tdir[i]), ctypes.byref(d), ctypes.byref(pp))
        tdist[i], tpdf[i] = d.value, pp.value
        parea[i] = lib.ref_projected_triangle_area(fp(tpos[i]), fp(tri[i, 0]), fp(tri[i, 1]), fp(tri[i, 2]))
    out.update(tri_verts=tri, tri_pos=tpos, tri_dir=tdir, tri_dist=tdist, tri_pdf=tpdf, tri_proj_area=parea)

    # Fresnel terms
    eta = rng.uniform(0.1, 3.0, size=n).astype(np.float32)
    kk = rng.uniform(0.0, 5.0, size=n).astype(np.float32)
    cosv = rng.uniform(0.0, 1.0, size=n).astype(np.float32)
    out.update(fresnel_eta=eta, fresnel_k=kk, fresnel_cos=cosv,
               fresnel_conductor=np.array([lib.ref_fresnel_conductor(F(a), F(b), F(c)) for a, b, c in zip(eta, kk, cosv)],
                                          dtype=np.float32),
               fresnel_schlick=np.array([lib.ref_fresnel_schlick(F(a / 3.0), F(c)) for a, c in zip(eta, cosv)],
                                        dtype=np.float32))
    eta_d = rng.uniform(1.0, 2.5, size=n).astype(np.float32)
    ct = np.array([lib.ref_cos_transmittance(F(a), F(c)) for a, c in zip(eta_d, cosv)], dtype=np.float32)
    out.update(dielectric_eta=eta_d, cos_transmittance=ct,
               fresnel_dielectric=np.array([lib.ref_fresnel_dielectric(F(a), F(c), F(t))
                                            for a, c, t in zip(eta_d, cosv, ct)], dtype=np.float32))

    # basis construction and reflection
    nv = rng.normal(size=(n, 3))
    nv /= np.linalg.norm(nv, axis=1, keepdims=True)
    nv = nv.astype(np.float32)
    nv[:6] = [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0], [0, -1, 0], [-1, 0, 0]]
    basis = np.zeros((n, 9), dtype=np.float32)
    refl = np.zeros((n, 3), dtype=np.float32)
    inc = rng.normal(size=(n, 3)).astype(np.float32)
    for i in range(n):
        lib.ref_make_basis(fp(nv[i]), fp(basis[i]))
        lib.ref_reflect(fp(inc[i]), fp(nv[i]), fp(refl[i]))
    out.update(basis_normal=nv, basis=basis, reflect_inc=inc, reflect=refl)
    np.savez_compressed(os.path.join(HERE, "ref_sampling.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_sampling.npz"), "with", len(out), "arrays")
    flake_fixture(lib)
    phong_fixture(lib)


def flake_fixture(lib):
    """tests/golden/ref_flake.npz: Matrix3 makeBasisTangent, transpose,
    product, transformCovariance, transformRotationInverse and Cholesky, and
    the SGGX micro-flake phase sample composed from them (ref_shim.cc)."""
    F = ctypes.c_float
    P = ctypes.POINTER(F)

    def fp(a):
        return a.ctypes.data_as(P)

    rng = np.random.Generator(np.random.PCG64(4242))
    n = 512
    out = {}
    tan = rng.normal(size=(n, 3))
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    tan = tan.astype(np.float32)
    tan[:6] = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]]
    bt = np.zeros((n, 9), dtype=np.float32)
    for i in range(n):
        lib.ref_make_basis_tangent(fp(tan[i]), fp(bt[i]))
    out.update(tangent=tan, basis_tangent=bt)
    a = rng.normal(size=(n, 9)).astype(np.float32)
    b = rng.normal(size=(n, 9)).astype(np.float32)
    var = rng.uniform(0.01, 2.0, size=(n, 3)).astype(np.float32)
    vec = rng.normal(size=(n, 3)).astype(np.float32)
    prod = np.zeros((n, 9), dtype=np.float32)
    tr = np.zeros((n, 9), dtype=np.float32)
    cov = np.zeros((n, 9), dtype=np.float32)
    rot_inv = np.zeros((n, 3), dtype=np.float32)
    for i in range(n):
        lib.ref_mat_mul(fp(a[i]), fp(b[i]), fp(prod[i]))
        lib.ref_transpose(fp(a[i]), fp(tr[i]))
        lib.ref_transform_covariance(fp(bt[i]), fp(var[i]), fp(cov[i]))
        lib.ref_transform_rotation_inverse(fp(a[i]), fp(vec[i]), fp(rot_inv[i]))
    # symmetric positive definite inputs for the Cholesky factor
    chol = np.zeros((n, 9), dtype=np.float32)
    for i in range(n):
        lib.ref_cholesky(fp(cov[i]), fp(chol[i]))
    out.update(mat_a=a, mat_b=b, mat_prod=prod, mat_transpose=tr, cov_var=var, cov=cov, rot_vec=vec,
               rot_inv=rot_inv, cholesky=chol)
    wo = rng.normal(size=(n, 3))
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wo = wo.astype(np.float32)
    sd = rng.uniform(0.05, 1.0, size=(n, 3)).astype(np.float32)
    sd[: n // 2] = [0.1, 1.0, 1.0]
    uv = rng.uniform(0.0, 1.0, size=(n, 2)).astype(np.float32)
    wi = np.zeros((n, 3), dtype=np.float32)
    for i in range(n):
        lib.ref_sggx_flake_sample(fp(wo[i]), fp(sd[i]), fp(bt[i]), F(uv[i, 0]), F(uv[i, 1]), fp(wi[i]))
    out.update(flake_wo=wo, flake_stddev=sd, flake_frame=bt, flake_uv=uv, flake_wi=wi)
    np.savez_compressed(os.path.join(HERE, "ref_flake.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_flake.npz"), "with", len(out), "arrays")


def phong_fixture(lib):
    """tests/golden/ref_phong.npz: PowerCosineSampleHemisphere
    (sampling3.hh:96-104), the half-vector warp of the Blinn-Phong sampler."""
    F = ctypes.c_float
    P = ctypes.POINTER(F)
    rng = np.random.Generator(np.random.PCG64(77))
    n = 512
    uv = rng.uniform(0.0, 1.0, size=(n, 2)).astype(np.float32)
    power = rng.choice(np.array([1.0, 4.0, 10.0, 48.0, 200.0], dtype=np.float32), size=n)
    res = np.zeros((n, 3), dtype=np.float32)
    for i in range(n):
        lib.ref_power_cosine_sample_hemisphere(F(uv[i, 0]), F(uv[i, 1]), F(power[i]), res[i].ctypes.data_as(P))
    np.savez_compressed(os.path.join(HERE, "ref_phong.npz"), uv=uv, power=power, power_cosine_sample_hemisphere=res)
    print("wrote", os.path.join(HERE, "ref_phong.npz"))


if __name__ == "__main__":
    main()
