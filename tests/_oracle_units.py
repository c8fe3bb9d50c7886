"""ctypes access to the oracle's unit exports (test infrastructure)."""
import ctypes

import numpy as np

import oracle
from tempestrenderer_amd import _abi

F = ctypes.c_float
PF = ctypes.POINTER(F)


def lib():
    lb = oracle.load()
    if not getattr(lb, "_units_declared", False):
        lb.orc_philox.argtypes = [ctypes.POINTER(ctypes.c_uint32)] * 3
        lb.orc_u01f.restype = F
        lb.orc_u01f.argtypes = [ctypes.c_uint32]
        lb.orc_fmath.argtypes = [ctypes.c_int, ctypes.c_size_t, PF, PF]
        lb.orc_bsdf_eval.argtypes = [ctypes.POINTER(_abi.tr_material), PF, PF, PF, PF, PF, F, PF]
        lb.orc_bsdf_sample.argtypes = [ctypes.POINTER(_abi.tr_material), PF, PF, PF, PF, PF, PF, F, PF]
        for name in ("orc_fresnel_conductor", "orc_fresnel_schlick", "orc_fresnel_dielectric"):
            getattr(lb, name).restype = F
            getattr(lb, name).argtypes = [F, F] if name == "orc_fresnel_schlick" else [F, F, F]
        lb.orc_make_basis.argtypes = [PF, PF]
        lb.orc_lcg15.argtypes = [ctypes.c_uint32, ctypes.c_uint32, PF]
        lb.orc_sample_sphere_light.argtypes = [PF, PF, F, F, F, PF]
        lb.orc_sphere_light_pdf.restype = F
        lb.orc_sphere_light_pdf.argtypes = [PF, PF, PF, F]
        lb.orc_sample_triangle_area.argtypes = [PF, PF, PF, PF, F, F, PF]
        lb.orc_projected_triangle_area.restype = F
        lb.orc_projected_triangle_area.argtypes = [PF, PF, PF, PF]
        lb.orc_sggx.argtypes = [PF, PF, PF, PF]
        lb.orc_make_basis_tangent.argtypes = [PF, PF]
        lb.orc_cholesky.argtypes = [PF, PF]
        lb.orc_transpose.argtypes = [PF, PF]
        lb.orc_mat_mul.argtypes = [PF, PF, PF]
        lb.orc_transform_covariance.argtypes = [PF, PF, PF]
        lb.orc_transform_rotation_inverse.argtypes = [PF, PF, PF]
        lb.orc_bsdf_eval_batch.argtypes = [ctypes.POINTER(_abi.tr_material), PF, PF, PF, PF, ctypes.c_size_t, PF, F,
                                           PF]
        lb._units_declared = True
    return lb


def fp(a):
    return a.ctypes.data_as(PF)


def f32(*v):
    return np.array(v, dtype=np.float32)


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox(c, k, o)
    return list(o)


def fmath(op, x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    lib().orc_fmath(op, x.size, fp(x), fp(out))
    return out


def bsdf_eval(mat, n, wo, wi, t=None, b=None, dir_density=-1.0):
    out = np.zeros(4, dtype=np.float32)
    tt = f32(*(t if t is not None else (0, 0, 0)))
    bb = f32(*(b if b is not None else (0, 0, 0)))
    lib().orc_bsdf_eval(ctypes.byref(mat.rec), fp(f32(*n)), fp(tt), fp(bb), fp(f32(*wo)), fp(f32(*wi)),
                        float(dir_density), fp(out))
    return out[:3], float(out[3])


def bsdf_sample(mat, n, wo, xi, strata=(0, 0, 1, 1), t=None, b=None, dir_density=0.0):
    """xi: three uniforms (cosine pair, lobe selector), optionally two more
    (the anisotropic microfacet models' specular pair, 0.5 by default)."""
    xi5 = tuple(xi) + (0.5, 0.5)[: 5 - len(xi)]
    out = np.zeros(4, dtype=np.float32)
    tt = f32(*(t if t is not None else (0, 0, 0)))
    bb = f32(*(b if b is not None else (0, 0, 0)))
    lib().orc_bsdf_sample(ctypes.byref(mat.rec), fp(f32(*n)), fp(tt), fp(bb), fp(f32(*wo)), fp(f32(*strata)),
                          fp(f32(*xi5)), float(dir_density), fp(out))
    return out[:3], float(out[3])
