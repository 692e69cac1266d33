// Device geometry core of the MCUT front end, in IEEE binary64.
//
// Every floating-point expression follows the reference's operation order
// (compiled with -ffp-contract=off, so no FMA contraction; fma() appears only
// inside the error-free product, where the exact rounding error is wanted).
// Division and square root lower to gfx950's correctly rounded sequences.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcgpu {

// ---------------------------------------------------------------------------
// Error-free transformations (Knuth two-sum, FMA two-product) and expansion
// arithmetic after Shewchuk (1997).  The error-bound coefficients are the
// reference's compile-time values for epsilon = 2^-53 (shewchuk.c:420-433).
// ---------------------------------------------------------------------------
constexpr double kResultErrBound = 3.3306690738754706e-16;
constexpr double kCcwErrBoundA = 3.3306690738754716e-16;
constexpr double kCcwErrBoundB = 2.2204460492503146e-16;
constexpr double kCcwErrBoundC = 1.1093356479670487e-31;
constexpr double kO3dErrBoundA = 7.7715611723761027e-16;

__device__ __forceinline__ void two_sum(double a, double b, double& x, double& y)
{
    x = a + b;
    double bv = x - a;
    double av = x - bv;
    y = (a - av) + (b - bv);
}

__device__ __forceinline__ void fast_two_sum(double a, double b, double& x, double& y)
{
    x = a + b;
    y = b - (x - a);
}

__device__ __forceinline__ void two_diff(double a, double b, double& x, double& y)
{
    x = a - b;
    double bv = a - x;
    double av = x + bv;
    y = (a - av) + (bv - b);
}

__device__ __forceinline__ double two_diff_tail(double a, double b, double x)
{
    double bv = a - x;
    double av = x + bv;
    return (a - av) + (bv - b);
}

__device__ __forceinline__ void two_product(double a, double b, double& x, double& y)
{
    x = a * b;
    y = __fma_rn(a, b, -x); // the exact rounding error of the product
}

// (a1 + a0) - (b1 + b0) -> 4 components, increasing magnitude
__device__ __forceinline__ void two_two_diff(double a1, double a0, double b1, double b0, double* x)
{
    double i, j, z0, i2;
    two_diff(a0, b0, i, x[0]);
    two_sum(a1, i, j, z0);
    two_diff(z0, b1, i2, x[1]);
    two_sum(j, i2, x[3], x[2]);
}

__device__ inline int fast_expansion_sum_zeroelim(int elen, const double* e, int flen, const double* f, double* h)
{
    double q, qn, hh, en = e[0], fn = f[0];
    int ei = 0, fi = 0, hi = 0;
    if ((fn > en) == (fn > -en)) {
        q = en;
        ++ei;
        en = ei < elen ? e[ei] : 0.0;
    } else {
        q = fn;
        ++fi;
        fn = fi < flen ? f[fi] : 0.0;
    }
    if (ei < elen && fi < flen) {
        if ((fn > en) == (fn > -en)) {
            fast_two_sum(en, q, qn, hh);
            ++ei;
            en = ei < elen ? e[ei] : 0.0;
        } else {
            fast_two_sum(fn, q, qn, hh);
            ++fi;
            fn = fi < flen ? f[fi] : 0.0;
        }
        q = qn;
        if (hh != 0.0) h[hi++] = hh;
        while (ei < elen && fi < flen) {
            if ((fn > en) == (fn > -en)) {
                two_sum(q, en, qn, hh);
                ++ei;
                en = ei < elen ? e[ei] : 0.0;
            } else {
                two_sum(q, fn, qn, hh);
                ++fi;
                fn = fi < flen ? f[fi] : 0.0;
            }
            q = qn;
            if (hh != 0.0) h[hi++] = hh;
        }
    }
    while (ei < elen) {
        two_sum(q, en, qn, hh);
        ++ei;
        en = ei < elen ? e[ei] : 0.0;
        q = qn;
        if (hh != 0.0) h[hi++] = hh;
    }
    while (fi < flen) {
        two_sum(q, fn, qn, hh);
        ++fi;
        fn = fi < flen ? f[fi] : 0.0;
        q = qn;
        if (hh != 0.0) h[hi++] = hh;
    }
    if (q != 0.0 || hi == 0) h[hi++] = q;
    return hi;
}

// Adaptive orient2d.  Its value (not only its sign) is used to rank vertex
// triplets (math.cpp:359-386), so the stages and their acceptance tests follow
// the published algorithm exactly (shewchuk.c:1611-1733).
__device__ inline double orient2d_adaptive(const double* pa, const double* pb, const double* pc)
{
    double detleft = (pa[0] - pc[0]) * (pb[1] - pc[1]);
    double detright = (pa[1] - pc[1]) * (pb[0] - pc[0]);
    double det = detleft - detright, detsum;
    if (detleft > 0.0) {
        if (detright <= 0.0) return det;
        detsum = detleft + detright;
    } else if (detleft < 0.0) {
        if (detright >= 0.0) return det;
        detsum = -detleft - detright;
    } else {
        return det;
    }
    double errbound = kCcwErrBoundA * detsum;
    if (det >= errbound || -det >= errbound) return det;

    double acx = pa[0] - pc[0], bcx = pb[0] - pc[0], acy = pa[1] - pc[1], bcy = pb[1] - pc[1];
    double dl, dlt, dr, drt, B[4], u[4], C1[8], C2[12], D[16];
    two_product(acx, bcy, dl, dlt);
    two_product(acy, bcx, dr, drt);
    two_two_diff(dl, dlt, dr, drt, B);
    det = B[0];
    for (int i = 1; i < 4; ++i) det += B[i];
    errbound = kCcwErrBoundB * detsum;
    if (det >= errbound || -det >= errbound) return det;

    double acxt = two_diff_tail(pa[0], pc[0], acx);
    double bcxt = two_diff_tail(pb[0], pc[0], bcx);
    double acyt = two_diff_tail(pa[1], pc[1], acy);
    double bcyt = two_diff_tail(pb[1], pc[1], bcy);
    if (acxt == 0.0 && acyt == 0.0 && bcxt == 0.0 && bcyt == 0.0) return det;

    errbound = kCcwErrBoundC * detsum + kResultErrBound * fabs(det);
    det += (acx * bcyt + bcy * acxt) - (acy * bcxt + bcx * acyt);
    if (det >= errbound || -det >= errbound) return det;

    double s1, s0, t1, t0;
    two_product(acxt, bcy, s1, s0);
    two_product(acyt, bcx, t1, t0);
    two_two_diff(s1, s0, t1, t0, u);
    int c1 = fast_expansion_sum_zeroelim(4, B, 4, u, C1);
    two_product(acx, bcyt, s1, s0);
    two_product(acy, bcxt, t1, t0);
    two_two_diff(s1, s0, t1, t0, u);
    int c2 = fast_expansion_sum_zeroelim(c1, C1, 4, u, C2);
    two_product(acxt, bcyt, s1, s0);
    two_product(acyt, bcxt, t1, t0);
    two_two_diff(s1, s0, t1, t0, u);
    int dn = fast_expansion_sum_zeroelim(c2, C2, 4, u, D);
    return D[dn - 1];
}

// orient3d semi-static filter (shewchuk.c:2367-2410, stage A).  Returns +1/-1
// when the floating-point determinant's sign is certain, 0 when it is not.
__device__ __forceinline__ int orient3d_filter(const double* pa, const double* pb, const double* pc, const double* pd)
{
    double adx = pa[0] - pd[0], bdx = pb[0] - pd[0], cdx = pc[0] - pd[0];
    double ady = pa[1] - pd[1], bdy = pb[1] - pd[1], cdy = pc[1] - pd[1];
    double adz = pa[2] - pd[2], bdz = pb[2] - pd[2], cdz = pc[2] - pd[2];
    double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    double cdxady = cdx * ady, adxcdy = adx * cdy;
    double adxbdy = adx * bdy, bdxady = bdx * ady;
    double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                  (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    double eb = kO3dErrBoundA * perm;
    if (det > eb) return 1;
    if (-det > eb) return -1;
    return 0;
}

// orient3d pre-filter in single precision, ahead of the FP64 filter.  The
// nine coordinate differences are formed in double, as the FP64 filter forms
// them, and rounded to float: each is within 2^-24 (1 + 2^-29) of the exact
// difference, one float rounding plus the double one.  The determinant and
// its permanent are then evaluated in float in the order of orient3dfast.
// Shewchuk's stage-A analysis bounds the error of that evaluation by
// (7 + 56 eps) eps * permanent for entries each within eps of exact; the
// bound used here is 16 eps * permanent with eps = 2^-24, which covers the
// slightly larger entry error with a factor of two to spare.  The analysis is
// relative, so it needs every product to be a normal float: each difference
// must be 0 or have magnitude in [2^-40, 2^40] (products of three then lie in
// [2^-120, 2^120]; a product difference that cancels to a subnormal has an
// absolute error of at most a few 2^-149, below the bound whenever the
// permanent is nonzero).  Outside that range, or for a NaN, the pre-filter
// declines.  Returns +1/-1 when the sign is certain, 0 when it is not.
constexpr float kO3dErrBoundF32 = 9.5367431640625e-07f; // 16 * 2^-24

__device__ __forceinline__ bool f32_in_range(float x)
{
    const float a = fabsf(x);
    return a == 0.0f || (a >= 9.094947017729282e-13f && a <= 1.099511627776e12f); // [2^-40, 2^40]
}

__device__ __forceinline__ int orient3d_filter_f32(const double* pa, const double* pb, const double* pc, const double* pd)
{
    const float adx = (float)(pa[0] - pd[0]), bdx = (float)(pb[0] - pd[0]), cdx = (float)(pc[0] - pd[0]);
    const float ady = (float)(pa[1] - pd[1]), bdy = (float)(pb[1] - pd[1]), cdy = (float)(pc[1] - pd[1]);
    const float adz = (float)(pa[2] - pd[2]), bdz = (float)(pb[2] - pd[2]), cdz = (float)(pc[2] - pd[2]);
    if (!(f32_in_range(adx) && f32_in_range(bdx) && f32_in_range(cdx) && f32_in_range(ady) && f32_in_range(bdy) &&
          f32_in_range(cdy) && f32_in_range(adz) && f32_in_range(bdz) && f32_in_range(cdz)))
        return 0; // NaN fails the range test too
    const float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const float cdxady = cdx * ady, adxcdy = adx * cdy;
    const float adxbdy = adx * bdy, bdxady = bdx * ady;
    const float det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    const float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * fabsf(adz) + (fabsf(cdxady) + fabsf(adxcdy)) * fabsf(bdz) +
                       (fabsf(adxbdy) + fabsf(bdxady)) * fabsf(cdz);
    const float eb = kO3dErrBoundF32 * perm;
    if (det > eb) return 1;
    if (-det > eb) return -1;
    return 0;
}

// the staged sign: the FP32 pre-filter, then the FP64 filter; 0 = undecided
// (escalate).  stage receives 0 (FP32 decided), 1 (FP64 decided) or 2.
__device__ __forceinline__ int orient3d_staged(const double* pa, const double* pb, const double* pc, const double* pd,
                                               int& stage)
{
    int s = orient3d_filter_f32(pa, pb, pc, pd);
    if (s != 0) {
        stage = 0;
        return s;
    }
    s = orient3d_filter(pa, pb, pc, pd);
    stage = s != 0 ? 1 : 2;
    return s;
}

// --- exact orient3d sign: expansions of the coordinate differences ----------

__device__ inline int grow_zeroelim(int n, const double* e, double b, double* h)
{
    double q = b, qn, hh;
    int hi = 0;
    for (int i = 0; i < n; ++i) {
        two_sum(q, e[i], qn, hh);
        q = qn;
        if (hh != 0.0) h[hi++] = hh;
    }
    if (q != 0.0 || hi == 0) h[hi++] = q;
    return hi;
}

__device__ inline int scale_zeroelim(int n, const double* e, double b, double* h)
{
    double q, hh, p1, p0, sum;
    int hi = 0;
    two_product(e[0], b, q, hh);
    if (hh != 0.0) h[hi++] = hh;
    for (int i = 1; i < n; ++i) {
        two_product(e[i], b, p1, p0);
        two_sum(q, p0, sum, hh);
        if (hh != 0.0) h[hi++] = hh;
        fast_two_sum(p1, sum, q, hh);
        if (hh != 0.0) h[hi++] = hh;
    }
    if (q != 0.0 || hi == 0) h[hi++] = q;
    return hi;
}

// h = e + f (in place into h; tmp is scratch of the same capacity)
__device__ inline int add_into(int n, double* h, int m, const double* f, double* tmp)
{
    for (int i = 0; i < m; ++i) {
        int l = grow_zeroelim(n, h, f[i], tmp);
        for (int k = 0; k < l; ++k) h[k] = tmp[k];
        n = l;
    }
    return n;
}

__device__ inline int diff_expansion(double a, double b, double* e)
{
    double x, y;
    two_diff(a, b, x, y);
    if (y != 0.0) {
        e[0] = y;
        e[1] = x;
        return 2;
    }
    e[0] = x;
    return 1;
}

// exact sign of det[a-d; b-d; c-d]
__device__ inline int orient3d_exact_sign(const double* pa, const double* pb, const double* pc, const double* pd)
{
    double X[3][2], Y[3][2], Z[3][2];
    int nx[3], ny[3], nz[3];
    const double* P[3] = {pa, pb, pc};
    for (int i = 0; i < 3; ++i) {
        nx[i] = diff_expansion(P[i][0], pd[0], X[i]);
        ny[i] = diff_expansion(P[i][1], pd[1], Y[i]);
        nz[i] = diff_expansion(P[i][2], pd[2], Z[i]);
    }
    double acc[200], tmp[200], prod[16], t1[8], t2[8], part[8], m[64];
    int alen = 1;
    acc[0] = 0.0;
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (i + 2) % 3;
        // t1 = X_j * Y_k
        int l1 = 1;
        t1[0] = 0.0;
        for (int q = 0; q < ny[k]; ++q) {
            int pl = scale_zeroelim(nx[j], X[j], Y[k][q], part);
            l1 = add_into(l1, t1, pl, part, tmp);
        }
        // t2 = -(X_k * Y_j)
        int l2 = 1;
        t2[0] = 0.0;
        for (int q = 0; q < ny[j]; ++q) {
            int pl = scale_zeroelim(nx[k], X[k], -Y[j][q], part);
            l2 = add_into(l2, t2, pl, part, tmp);
        }
        for (int q = 0; q < l1; ++q) prod[q] = t1[q];
        int lm = add_into(l1, prod, l2, t2, tmp);
        // m = prod * Z_i
        int ml = 1;
        m[0] = 0.0;
        for (int q = 0; q < nz[i]; ++q) {
            double sc[40];
            int pl = scale_zeroelim(lm, prod, Z[i][q], sc);
            ml = add_into(ml, m, pl, sc, tmp);
        }
        alen = add_into(alen, acc, ml, m, tmp);
    }
    double top = acc[alen - 1];
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

// ---------------------------------------------------------------------------
// Geometry of math.cpp in the reference's operation order
// ---------------------------------------------------------------------------

__device__ __forceinline__ double dot3(const double* a, const double* b)
{
    double s = 0.0; // dot_product starts from 0.0 (math.h:634-642)
    s += a[0] * b[0];
    s += a[1] * b[1];
    s += a[2] * b[2];
    return s;
}

// compute_polygon_plane_coefficients (math.cpp:130-239).  Returns the largest
// normal component, or -1 for a degenerate polygon (normal set to zero).
template <typename VertexAt>
__device__ inline int plane_coefficients(int n, VertexAt vat, double* normal, double* d)
{
    double nx = 0.0, ny = 0.0, nz = 0.0;
    double c[3], x[3], p0[3];
    vat(0, p0);
    for (int i = 0; i < n; ++i) {
        vat(i, c);
        vat(i + 1 == n ? 0 : i + 1, x);
        nx += (c[1] - x[1]) * (c[2] + x[2]);
        ny += (c[2] - x[2]) * (c[0] + x[0]);
        nz += (c[0] - x[0]) * (c[1] + x[1]);
    }
    double nn[3] = {nx, ny, nz};
    double sl = dot3(nn, nn);
    if (isnan(nx) || isnan(ny) || isnan(nz) || sl < 1e-9 || sl == 0.0) {
        normal[0] = normal[1] = normal[2] = 0.0;
        *d = 0.0;
        return -1;
    }
    double len = sqrt(sl);
    normal[0] = nx / len;
    normal[1] = ny / len;
    normal[2] = nz / len;
    *d = dot3(p0, normal);
    double largest = 0.0;
    int k = 0;
    for (int i = 0; i < 3; ++i) {
        double a = fabs(normal[i]);
        if (a > largest) {
            largest = a;
            k = i;
        }
    }
    return k;
}

// calculate_projection_matrix (math.cpp:710-793) times the 2x3 selector:
// the two rows of K*R (row-major 2x3 in M)
__device__ inline void projection_matrix(const double* normal, int k, double* M)
{
    double l = sqrt(dot3(normal, normal));
    double a[3] = {normal[0] / l, normal[1] / l, normal[2] / l};
    double nk = normal[k];
    double s = (nk > 0.0) ? 1.0 : ((nk < 0.0) ? -1.0 : 0.0);
    double b[3] = {0.0, 0.0, 0.0};
    b[k] = 1.0 * s;
    double R[9] = {1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, 1.0};
    if (a[0] != b[0] || a[1] != b[1] || a[2] != b[2]) {
        double u[3] = {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
        double ab = dot3(a, b);
        const double I[9] = {1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, 1.0};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) R[3 * r + c] = ((u[r] * u[c]) / ab) * 2.0 - I[3 * r + c];
    }
    // rows of the selector: k=0 -> (y, z), k=1 -> (x, z), k=2 -> (x, y)
    const int r0 = (k == 0) ? 1 : 0;
    const int r1 = (k == 2) ? 1 : 2;
    for (int c = 0; c < 3; ++c) {
        // matrix product accumulates (0 + 0*R0c) + 1*R1c ... in index order (math.h:450-463)
        double acc0 = 0.0, acc1 = 0.0;
        for (int q = 0; q < 3; ++q) {
            acc0 += (q == r0 ? 1.0 : 0.0) * R[3 * q + c];
            acc1 += (q == r1 ? 1.0 : 0.0) * R[3 * q + c];
        }
        M[c] = acc0;
        M[3 + c] = acc1;
    }
}

// matrix_t * vec3 (math.h:508-531)
__device__ __forceinline__ void project2(const double* M, const double* x, double* out)
{
    double r0 = 0.0, r1 = 0.0;
    for (int c = 0; c < 3; ++c) {
        r0 = r0 + (M[c] * x[c]);
        r1 = r1 + (M[3 + c] * x[c]);
    }
    out[0] = r0;
    out[1] = r1;
}

// compute_point_in_polygon_test (math.cpp:851-902 -> 645-703)
template <typename VertexAt>
__device__ inline char point_in_polygon(const double* p, int n, VertexAt vat, const double* normal, int k)
{
    double M[6], pp[2];
    projection_matrix(normal, k, M);
    project2(M, p, pp);
    int rc = 0, lc = 0;
    double vi[3], w_il[2], w_i[2], first[2];
    // vertex n-1 is the predecessor of vertex 0
    vat(n - 1, vi);
    project2(M, vi, w_il);
    w_il[0] = w_il[0] - pp[0];
    w_il[1] = w_il[1] - pp[1];
    (void)first;
    for (int i = 0; i < n; ++i) {
        vat(i, vi);
        project2(M, vi, w_i);
        w_i[0] = w_i[0] - pp[0];
        w_i[1] = w_i[1] - pp[1];
        if (w_i[0] == 0.0 && w_i[1] == 0.0) return 'v';
        bool rs = (w_i[1] > 0.0) != (w_il[1] > 0.0);
        bool ls = (w_i[1] < 0.0) != (w_il[1] < 0.0);
        if (rs || ls) {
            double x = (w_i[0] * w_il[1] - w_il[0] * w_i[1]) / (w_il[1] - w_i[1]);
            if (rs && x > 0.0) rc++;
            if (ls && x < 0.0) lc++;
        }
        w_il[0] = w_i[0];
        w_il[1] = w_i[1];
    }
    if ((rc % 2) != (lc % 2)) return 'e';
    return (rc % 2) == 1 ? 'i' : 'o';
}

// compute_segment_plane_intersection (math.cpp:249-287); p untouched when the
// denominator is zero (the caller starts from (0,0,0), kernel.cpp:2505)
__device__ __forceinline__ void segment_point(const double* normal, double d, const double* q, const double* r, double* p)
{
    double num = d - dot3(q, normal);
    double rq[3] = {r[0] - q[0], r[1] - q[1], r[2] - q[2]};
    double den = dot3(rq, normal);
    if (den == 0.0) return;
    double t = num / den;
    for (int i = 0; i < 3; ++i) p[i] = q[i] + t * (r[i] - q[i]);
}

__device__ __forceinline__ char segment_type(int sq, int sr)
{
    if (sq == 0 && sr == 0) return 'p';
    if (sq == 0) return 'q';
    if (sr == 0) return 'r';
    if (sq == sr) return '0';
    return '1';
}

} // namespace mcgpu
