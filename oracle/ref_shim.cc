/*
 * ref_shim.cc - C exports around the reference's own header-only sampling
 * code, compiled from the sources under the reference checkout into
 * oracle/_ref/libtempest_ref.so by oracle/Makefile (target "ref").
 *
 * ORACLE: test infrastructure.  It pins the oracle's restatement of the
 * light and hemisphere warps, Fresnel terms and basis construction against
 * the reference functions themselves:
 *   include/tempest/math/sampling3.hh:55-264 (warps, sphere/triangle area)
 *   include/tempest/math/functions.hh:262-301 (Fresnel)
 *   include/tempest/math/matrix3.hh:92-191 (makeBasis, makeBasisTangent,
 *     transform, transformCovariance, transpose, product, Cholesky)
 *   include/tempest/math/vector3.hh:420 (Reflect)
 * Only headers that compile without the generated tempest/utils/config.hh
 * are used.  The reference computes sin/cos with the tan-half-angle
 * FastSinCos (functions.hh:322-353); comparisons allow for that.
 */
#include "tempest/math/functions.hh"
#include "tempest/math/vector3.hh"
#include "tempest/math/matrix3.hh"
#include "tempest/math/sampling3.hh"

using namespace Tempest;

static Vector3 V(const float* f) { return Vector3{f[0], f[1], f[2]}; }
static void W(const Vector3& v, float* f) { f[0] = v.x; f[1] = v.y; f[2] = v.z; }

extern "C" {

void ref_cosine_sample_hemisphere(float ra, float rb, float* out) { W(CosineSampleHemisphere(ra, rb), out); }
void ref_power_cosine_sample_hemisphere(float ra, float rb, float power, float* out)
{
    W(PowerCosineSampleHemisphere(ra, rb, power), out);
}
void ref_uniform_sample_hemisphere(float ra, float rb, float* out) { W(UniformSampleHemisphere(ra, rb), out); }
void ref_uniform_sample_sphere(float ra, float rb, float* out) { W(UniformSampleSphere(ra, rb), out); }

void ref_sample_sphere_area(const float* pos, const float* c, float r, float ra, float rb, float* dir,
                            float* dist, float* pdf)
{
    SampleIntersect s = SampleSphereArea(V(pos), V(c), r, ra, rb);
    W(s.Direction, dir);
    *dist = s.Distance;
    *pdf = s.PDF;
}

float ref_sphere_area_pdf(const float* pos, const float* dir, const float* c, float r)
{
    return ComputeSphereAreaPDF(V(pos), V(dir), V(c), r);
}

void ref_sample_triangle_area(const float* pos, const float* p0, const float* p1, const float* p2, float ra,
                              float rb, float* dir, float* dist, float* pdf)
{
    SampleIntersect s = UniformSampleTriangleArea(V(pos), V(p0), V(p1), V(p2), ra, rb);
    W(s.Direction, dir);
    *dist = s.Distance;
    *pdf = s.PDF;
}

float ref_projected_triangle_area(const float* pos, const float* p0, const float* p1, const float* p2)
{
    return ProjectedTriangleSphereAreaApproximate(V(pos), V(p0), V(p1), V(p2));
}

float ref_fresnel_conductor(float eta, float k, float c) { return FresnelConductor(eta, k, c); }
float ref_fresnel_schlick(float r, float c) { return FresnelSchlick(r, c); }
float ref_fresnel_dielectric(float eta, float ci, float ct) { return FresnelDielectric(eta, ci, ct); }
float ref_cos_transmittance(float eta, float ci) { return CosTransmittance(eta, ci); }

void ref_make_basis(const float* n, float* out9)
{
    Matrix3 m;
    m.makeBasis(V(n));
    W(m.column(0), out9);
    W(m.column(1), out9 + 3);
    W(m.column(2), out9 + 6);
}

void ref_basis_transform(const float* n, const float* v, float* out)
{
    Matrix3 m;
    m.makeBasis(V(n));
    W(m.transform(V(v)), out);
}

void ref_reflect(const float* inc, const float* n, float* out) { W(Reflect(V(inc), V(n)), out); }

float ref_specular_power_to_alpha2(float p) { return 2.0f / (p + 2.0f); }

/* Matrix3 pieces of the SGGX micro-flake phase sampling (matrix3.hh:92-191):
 * matrices pass as 9 floats, column by column. */
static Matrix3 M(const float* f) { return Matrix3(V(f), V(f + 3), V(f + 6)); }
static void WM(const Matrix3& m, float* f) { W(m.column(0), f); W(m.column(1), f + 3); W(m.column(2), f + 6); }

void ref_make_basis_tangent(const float* t, float* out9)
{
    Matrix3 m;
    m.makeBasisTangent(V(t));
    WM(m, out9);
}
void ref_cholesky(const float* m9, float* out9) { WM(M(m9).choleskyDecomposition(), out9); }
void ref_transpose(const float* m9, float* out9) { Matrix3 m = M(m9); WM(m.transpose(), out9); }
void ref_mat_mul(const float* a9, const float* b9, float* out9) { WM(M(a9) * M(b9), out9); }
void ref_transform_covariance(const float* m9, const float* var, float* out9)
{
    WM(M(m9).transformCovariance(V(var)), out9);
}
void ref_transform_rotation_inverse(const float* m9, const float* v, float* out)
{
    W(M(m9).transformRotationInverse(V(v)), out);
}

/* The body of SGGXMicroFlakeSampleIncidentLight (ggx-models-impl.hh:
 * 554-570), whose header needs the generated config.hh, written out over
 * the reference's own compiled primitives; `frame9` is the (tangent,
 * binormal, normal) basis. */
void ref_sggx_flake_sample(const float* wo, const float* stddev, const float* frame9, float ra, float rb, float* out)
{
    Vector3 uniform_sample = UniformSampleSphere(ra, rb);
    Vector3 sd = V(stddev);
    Matrix3 S = M(frame9).transformCovariance(sd * sd);
    Matrix3 out_basis;
    out_basis.makeBasis(V(wo));
    S = out_basis * S * out_basis.transpose();
    Vector3 vn = Normalize(S.choleskyDecomposition().transformRotationInverse(uniform_sample));
    vn = out_basis.transform(vn);
    W(Reflect(V(wo), vn), out);
}

}
