/* CPU oracle for the MCUT intersection front end -- TEST INFRASTRUCTURE ONLY.
 *
 * A single-threaded plain-C restatement of MCUT's front end: vertex shift,
 * halfedge numbering of the polygon soup, face AABBs, candidate face pairs,
 * edge-to-face test expansion and culling, face planes, the edge-face narrow
 * phase and the intersection registry with its status rules.  It exists to
 * check the HIP implementation; the product library never links it.
 *
 * Arithmetic follows the reference operation by operation (IEEE binary64,
 * round to nearest even, no contraction: build with -ffp-contract=off).
 * Reference citations are file:line in the MCUT source tree.
 */
#define _POSIX_C_SOURCE 199309L
#include "frontend_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ======================================================================
 * Exact arithmetic.  Error-free transformations and expansion arithmetic
 * after J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
 * Robust Geometric Predicates" (1997).  The reference compiles Shewchuk's
 * predicates with error-bound coefficients fixed at epsilon = 2^-53
 * (shewchuk.c:420-433); the same values are used here because orient2d's
 * returned value (not only its sign) depends on which stage accepts it.
 * ====================================================================== */
static const double k_resulterrbound = 3.3306690738754706e-16; /* (3 + 8e)e   */
static const double k_ccwerrboundA = 3.3306690738754716e-16;   /* (3 + 16e)e  */
static const double k_ccwerrboundB = 2.2204460492503146e-16;   /* (2 + 12e)e  */
static const double k_ccwerrboundC = 1.1093356479670487e-31;   /* (9 + 64e)e^2 */
static const double k_o3derrboundA = 7.7715611723761027e-16;   /* (7 + 56e)e  */

static inline void two_sum(double a, double b, double* x, double* y)
{
    double s = a + b, bv = s - a, av = s - bv;
    *x = s;
    *y = (a - av) + (b - bv);
}

static inline void fast_two_sum(double a, double b, double* x, double* y)
{
    double s = a + b;
    *x = s;
    *y = b - (s - a);
}

static inline void two_diff(double a, double b, double* x, double* y)
{
    double s = a - b, bv = a - s, av = s + bv;
    *x = s;
    *y = (a - av) + (bv - b);
}

static inline double two_diff_tail(double a, double b, double x)
{
    double bv = a - x, av = x + bv;
    return (a - av) + (bv - b);
}

static inline void two_product(double a, double b, double* x, double* y)
{
    double p = a * b;
    *x = p;
    *y = fma(a, b, -p); /* exact product error (unique for p = fl(a*b)) */
}

/* (a1 + a0) - (b1 + b0) as a 4-component expansion x[0..3] (increasing magnitude) */
static void two_two_diff(double a1, double a0, double b1, double b0, double* x)
{
    double i, j, z0, x0;
    /* Two_One_Diff(a1, a0, b0 -> j, z0, x[0]) */
    two_diff(a0, b0, &i, &x0);
    x[0] = x0;
    two_sum(a1, i, &j, &z0);
    /* Two_One_Diff(j, z0, b1 -> x3, x2, x1) */
    double i2, x1, x2, x3;
    two_diff(z0, b1, &i2, &x1);
    two_sum(j, i2, &x3, &x2);
    x[1] = x1;
    x[2] = x2;
    x[3] = x3;
}

static double estimate(int n, const double* e)
{
    double q = e[0];
    for (int i = 1; i < n; ++i) q += e[i];
    return q;
}

/* merge-sum of two strongly nonoverlapping expansions with zero elimination */
static int fast_expansion_sum_zeroelim(int elen, const double* e, int flen, const double* f, double* h)
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
            fast_two_sum(en, q, &qn, &hh);
            ++ei;
            en = ei < elen ? e[ei] : 0.0;
        } else {
            fast_two_sum(fn, q, &qn, &hh);
            ++fi;
            fn = fi < flen ? f[fi] : 0.0;
        }
        q = qn;
        if (hh != 0.0) h[hi++] = hh;
        while (ei < elen && fi < flen) {
            if ((fn > en) == (fn > -en)) {
                two_sum(q, en, &qn, &hh);
                ++ei;
                en = ei < elen ? e[ei] : 0.0;
            } else {
                two_sum(q, fn, &qn, &hh);
                ++fi;
                fn = fi < flen ? f[fi] : 0.0;
            }
            q = qn;
            if (hh != 0.0) h[hi++] = hh;
        }
    }
    while (ei < elen) {
        two_sum(q, en, &qn, &hh);
        ++ei;
        en = ei < elen ? e[ei] : 0.0;
        q = qn;
        if (hh != 0.0) h[hi++] = hh;
    }
    while (fi < flen) {
        two_sum(q, fn, &qn, &hh);
        ++fi;
        fn = fi < flen ? f[fi] : 0.0;
        q = qn;
        if (hh != 0.0) h[hi++] = hh;
    }
    if (q != 0.0 || hi == 0) h[hi++] = q;
    return hi;
}

/* adaptive orient2d; its returned value follows the published staging
 * (shewchuk.c:1611-1733) so that magnitudes compare equal to the reference's */
double om_orient2d(const double* pa, const double* pb, const double* pc)
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
    double errbound = k_ccwerrboundA * detsum;
    if (det >= errbound || -det >= errbound) return det;

    /* adaptive stages */
    double acx = pa[0] - pc[0], bcx = pb[0] - pc[0], acy = pa[1] - pc[1], bcy = pb[1] - pc[1];
    double dl, dlt, dr, drt, B[4], u[4], C1[8], C2[12], D[16];
    two_product(acx, bcy, &dl, &dlt);
    two_product(acy, bcx, &dr, &drt);
    two_two_diff(dl, dlt, dr, drt, B);
    det = estimate(4, B);
    errbound = k_ccwerrboundB * detsum;
    if (det >= errbound || -det >= errbound) return det;

    double acxt = two_diff_tail(pa[0], pc[0], acx);
    double bcxt = two_diff_tail(pb[0], pc[0], bcx);
    double acyt = two_diff_tail(pa[1], pc[1], acy);
    double bcyt = two_diff_tail(pb[1], pc[1], bcy);
    if (acxt == 0.0 && acyt == 0.0 && bcxt == 0.0 && bcyt == 0.0) return det;

    errbound = k_ccwerrboundC * detsum + k_resulterrbound * fabs(det);
    det += (acx * bcyt + bcy * acxt) - (acy * bcxt + bcx * acyt);
    if (det >= errbound || -det >= errbound) return det;

    double s1, s0, t1, t0;
    two_product(acxt, bcy, &s1, &s0);
    two_product(acyt, bcx, &t1, &t0);
    two_two_diff(s1, s0, t1, t0, u);
    int c1 = fast_expansion_sum_zeroelim(4, B, 4, u, C1);
    two_product(acx, bcyt, &s1, &s0);
    two_product(acy, bcxt, &t1, &t0);
    two_two_diff(s1, s0, t1, t0, u);
    int c2 = fast_expansion_sum_zeroelim(c1, C1, 4, u, C2);
    two_product(acxt, bcyt, &s1, &s0);
    two_product(acyt, bcxt, &t1, &t0);
    two_two_diff(s1, s0, t1, t0, u);
    int dn = fast_expansion_sum_zeroelim(c2, C2, 4, u, D);
    return D[dn - 1];
}

/* grow an expansion by one double (valid for any nonoverlapping input) */
static int grow_expansion_zeroelim(int n, const double* e, double b, double* h)
{
    double q = b, qn, hh;
    int hi = 0;
    for (int i = 0; i < n; ++i) {
        two_sum(q, e[i], &qn, &hh);
        q = qn;
        if (hh != 0.0) h[hi++] = hh;
    }
    if (q != 0.0 || hi == 0) h[hi++] = q;
    return hi;
}

static int scale_expansion_zeroelim(int n, const double* e, double b, double* h)
{
    double q, hh, p1, p0, sum;
    int hi = 0;
    two_product(e[0], b, &q, &hh);
    if (hh != 0.0) h[hi++] = hh;
    for (int i = 1; i < n; ++i) {
        two_product(e[i], b, &p1, &p0);
        two_sum(q, p0, &sum, &hh);
        if (hh != 0.0) h[hi++] = hh;
        fast_two_sum(p1, sum, &q, &hh);
        if (hh != 0.0) h[hi++] = hh;
    }
    if (q != 0.0 || hi == 0) h[hi++] = q;
    return hi;
}

/* h = e + f, by repeated growth (simple and valid for any nonoverlapping inputs) */
static int expansion_add(int n, const double* e, int m, const double* f, double* h)
{
    double tmp[512];
    int len = n;
    memcpy(h, e, (size_t)n * sizeof(double));
    for (int i = 0; i < m; ++i) {
        len = grow_expansion_zeroelim(len, h, f[i], tmp);
        memcpy(h, tmp, (size_t)len * sizeof(double));
    }
    return len;
}

static int expansion_mul(int n, const double* e, int m, const double* f, double* h)
{
    double part[512], acc[512];
    int len = 1;
    acc[0] = 0.0;
    for (int i = 0; i < m; ++i) {
        int pl = scale_expansion_zeroelim(n, e, f[i], part);
        len = expansion_add(len, acc, pl, part, h);
        memcpy(acc, h, (size_t)len * sizeof(double));
    }
    memcpy(h, acc, (size_t)len * sizeof(double));
    return len;
}

static int diff_expansion(double a, double b, double* e)
{
    double x, y;
    two_diff(a, b, &x, &y);
    if (y != 0.0) {
        e[0] = y;
        e[1] = x;
        return 2;
    }
    e[0] = x;
    return 1;
}

/* exact sign of orient3d(a,b,c,d) = det[a-d; b-d; c-d] (shewchuk.c:2367-2410 gives the
 * same sign: stage A filter first, then an exact expansion evaluation). */
int om_orient3d_sign(const double* pa, const double* pb, const double* pc, const double* pd)
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
    double errbound = k_o3derrboundA * perm;
    if (det > errbound || -det > errbound) return det > 0 ? 1 : -1;

    /* exact: differences as expansions, then the cofactor expansion along z */
    double X[3][2], Y[3][2], Z[3][2];
    int nx[3], ny[3], nz[3];
    const double* P[3] = {pa, pb, pc};
    for (int i = 0; i < 3; ++i) {
        nx[i] = diff_expansion(P[i][0], pd[0], X[i]);
        ny[i] = diff_expansion(P[i][1], pd[1], Y[i]);
        nz[i] = diff_expansion(P[i][2], pd[2], Z[i]);
    }
    double acc[512], t1[64], t2[64], t3[512], tmp[512];
    int alen = 1;
    acc[0] = 0.0;
    /* det = sum_i z_i * (x_j*y_k - x_k*y_j) for (i,j,k) cyclic */
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (i + 2) % 3;
        int l1 = expansion_mul(nx[j], X[j], ny[k], Y[k], t1);
        int l2 = expansion_mul(nx[k], X[k], ny[j], Y[j], t2);
        for (int q = 0; q < l2; ++q) t2[q] = -t2[q];
        int lm = expansion_add(l1, t1, l2, t2, tmp);
        int l3 = expansion_mul(lm, tmp, nz[i], Z[i], t3);
        alen = expansion_add(alen, acc, l3, t3, tmp);
        memcpy(acc, tmp, (size_t)alen * sizeof(double));
    }
    double top = acc[alen - 1];
    return top > 0.0 ? 1 : (top < 0.0 ? -1 : 0);
}

/* ======================================================================
 * Geometry core (math.cpp / math.h)
 * ====================================================================== */

static inline double dot3(const double* a, const double* b)
{
    double s = 0.0; /* math.h:634-642: accumulate from 0.0 in index order */
    s += a[0] * b[0];
    s += a[1] * b[1];
    s += a[2] * b[2];
    return s;
}

/* compute_polygon_plane_coefficients, math.cpp:130-239 */
int om_plane(const double* P, int n, double* normal, double* d)
{
    double nx = 0.0, ny = 0.0, nz = 0.0;
    for (int i = 0; i < n; ++i) {
        int j = (i + 1) % n;
        const double* c = P + 3 * i;
        const double* x = P + 3 * j;
        nx += (c[1] - x[1]) * (c[2] + x[2]);
        ny += (c[2] - x[2]) * (c[0] + x[0]);
        nz += (c[0] - x[0]) * (c[1] + x[1]);
    }
    double nn[3] = {nx, ny, nz};
    if (isnan(nx) || isnan(ny) || isnan(nz) || dot3(nn, nn) < 1e-9) {
        normal[0] = normal[1] = normal[2] = 0.0;
        *d = 0.0;
        return 0;
    }
    double sl = dot3(nn, nn);
    if (sl == 0.0) {
        normal[0] = nx, normal[1] = ny, normal[2] = nz;
        *d = 0.0;
        return 0;
    }
    double len = sqrt(sl); /* normalize(): v / sqrt(dot(v,v)), math.h:717-721 */
    normal[0] = nx / len;
    normal[1] = ny / len;
    normal[2] = nz / len;
    *d = dot3(P, normal);
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

/* calculate_projection_matrix (math.cpp:710-793) followed by the 2x3 selector
 * product K*R; returns the 2x3 matrix row-major in M. */
static void projection_matrix(const double* normal, int k, double* M)
{
    double l = sqrt(dot3(normal, normal));
    double a[3] = {normal[0] / l, normal[1] / l, normal[2] / l};
    double s = (0.0 < normal[k]) - (normal[k] < 0.0) > 0 ? 1.0 : ((normal[k] < 0.0) ? -1.0 : 0.0);
    double b[3] = {0.0, 0.0, 0.0};
    b[k] = 1.0 * s;
    double I[9] = {1.0, 0.0, 0.0, 0.0, -1.0, 0.0, 0.0, 0.0, 1.0};
    double R[9];
    memcpy(R, I, sizeof(R));
    if (a[0] != b[0] || a[1] != b[1] || a[2] != b[2]) {
        double u[3] = {a[0] + b[0], a[1] + b[1], a[2] + b[2]};
        double ab = dot3(a, b);
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) {
                double o = u[r] * u[c]; /* outer_product: column c = u * u[c] */
                R[3 * r + c] = ((o / ab) * 2.0) - I[3 * r + c];
            }
    }
    double K[6] = {0};
    if (k == 0) {
        K[1] = 1.0;
        K[5] = 1.0;
    } else if (k == 1) {
        K[0] = 1.0;
        K[5] = 1.0;
    } else {
        K[0] = 1.0;
        K[4] = 1.0;
    }
    for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 3; ++c) {
            double acc = 0.0; /* matrix_t::operator*, math.h:450-463 */
            for (int q = 0; q < 3; ++q) acc += K[3 * r + q] * R[3 * q + c];
            M[3 * r + c] = acc;
        }
}

/* matrix_t::operator*(vec3), math.h:508-531 (column-major accumulation) */
static inline void project(const double* M, const double* x, double* out)
{
    double r0 = 0.0, r1 = 0.0;
    for (int c = 0; c < 3; ++c) {
        r0 = r0 + (M[c] * x[c]);
        r1 = r1 + (M[3 + c] * x[c]);
    }
    out[0] = r0;
    out[1] = r1;
}

/* 2-D crossing test, math.cpp:645-703 (O'Rourke) */
static char pip2d(const double* q, const double* V2, int n)
{
    int rc = 0, lc = 0;
    for (int i = 0; i < n; ++i) {
        /* w_i = V2[i] - q, each difference one rounding as the reference forms it */
        int il = (i + n - 1) % n;
        const double wi0 = V2[2 * i] - q[0], wi1 = V2[2 * i + 1] - q[1];
        const double wl0 = V2[2 * il] - q[0], wl1 = V2[2 * il + 1] - q[1];
        if (wi0 == 0.0 && wi1 == 0.0) return 'v';
        int rs = (wi1 > 0.0) != (wl1 > 0.0);
        int ls = (wi1 < 0.0) != (wl1 < 0.0);
        if (rs || ls) {
            double x = (wi0 * wl1 - wl0 * wi1) / (wl1 - wi1);
            if (rs && x > 0.0) rc++;
            if (ls && x < 0.0) lc++;
        }
    }
    if ((rc % 2) != (lc % 2)) return 'e';
    return (rc % 2) == 1 ? 'i' : 'o';
}

/* compute_point_in_polygon_test (3-D), math.cpp:851-902 */
char om_point_in_polygon(const double* p, const double* P, int n, const double* normal, int k)
{
    double M[6], pp[2];
    double* V2 = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    projection_matrix(normal, k, M);
    project(M, p, pp);
    for (int i = 0; i < n; ++i) project(M, P + 3 * i, V2 + 2 * i);
    const char res = pip2d(pp, V2, n);
    free(V2);
    return res;
}

/* determine_three_noncollinear_vertices, math.cpp:289-389: every (i<j<k)
 * triplet with non-zero orient2d of the projected polygon, in lexicographic
 * order, is sorted by std::sort with "greater |orient2d|" and the first element
 * is taken.  The winner among equal keys depends on the sort's element moves, so
 * the GCC libstdc++ std::sort is replayed below: introsort with a 16-element
 * insertion-sort threshold, median-of-three pivot moved to the front, unguarded
 * partition, heap sort of a range whose depth limit (2 floor(log2 n)) runs out,
 * and a final insertion sort.  Returns 0 when every triplet is collinear. */
typedef struct {
    int i, j, k;
    double a; /* |orient2d| */
} om_trip;

static inline int trip_less(const om_trip* x, const om_trip* y) { return x->a > y->a; }
static inline void trip_swap(om_trip* x, om_trip* y)
{
    om_trip t = *x;
    *x = *y;
    *y = t;
}

static void trip_linear_insert(om_trip* last)
{
    om_trip val = *last;
    om_trip* next = last - 1;
    while (trip_less(&val, next)) {
        *last = *next;
        last = next;
        --next;
    }
    *last = val;
}

static void trip_insertion_sort(om_trip* first, om_trip* last)
{
    if (first == last) return;
    for (om_trip* i = first + 1; i != last; ++i) {
        if (trip_less(i, first)) {
            om_trip val = *i;
            memmove(first + 1, first, (size_t)(i - first) * sizeof(om_trip));
            *first = val;
        } else {
            trip_linear_insert(i);
        }
    }
}

static void trip_median_to_first(om_trip* result, om_trip* a, om_trip* b, om_trip* c)
{
    if (trip_less(a, b)) {
        if (trip_less(b, c)) trip_swap(result, b);
        else if (trip_less(a, c)) trip_swap(result, c);
        else trip_swap(result, a);
    } else if (trip_less(a, c)) {
        trip_swap(result, a);
    } else if (trip_less(b, c)) {
        trip_swap(result, c);
    } else {
        trip_swap(result, b);
    }
}

static om_trip* trip_partition(om_trip* first, om_trip* last, om_trip* pivot)
{
    for (;;) {
        while (trip_less(first, pivot)) ++first;
        --last;
        while (trip_less(pivot, last)) --last;
        if (!(first < last)) return first;
        trip_swap(first, last);
        ++first;
    }
}

/* the binary heap of libstdc++'s heap sort (a max-heap under trip_less):
 * sift the hole at `hole` down to a leaf along the larger children, then
 * push `value` up from there, no higher than `top` */
static void trip_push_heap(om_trip* first, long hole, long top, om_trip value)
{
    long parent = (hole - 1) / 2;
    while (hole > top && trip_less(first + parent, &value)) {
        first[hole] = first[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    first[hole] = value;
}

static void trip_adjust_heap(om_trip* first, long hole, long len, om_trip value)
{
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (trip_less(first + child, first + (child - 1))) child--;
        first[hole] = first[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        first[hole] = first[child - 1];
        hole = child - 1;
    }
    trip_push_heap(first, hole, top, value);
}

/* std::__partial_sort(first, last, last): make the heap, then pop it */
static void trip_heap_sort(om_trip* first, om_trip* last)
{
    const long len = last - first;
    if (len >= 2)
        for (long parent = (len - 2) / 2;; --parent) {
            trip_adjust_heap(first, parent, len, first[parent]);
            if (parent == 0) break;
        }
    while (last - first > 1) {
        --last;
        om_trip value = *last;
        *last = *first;
        trip_adjust_heap(first, 0, last - first, value);
    }
}

static void trip_introsort_loop(om_trip* first, om_trip* last, int depth)
{
    while (last - first > 16) {
        if (depth == 0) {
            trip_heap_sort(first, last);
            return;
        }
        --depth;
        om_trip* mid = first + (last - first) / 2;
        trip_median_to_first(first, first + 1, mid, last - 1);
        om_trip* cut = trip_partition(first + 1, last, first);
        trip_introsort_loop(cut, last, depth);
        last = cut;
    }
}

static void trip_sort(om_trip* first, om_trip* last)
{
    long n = last - first;
    if (n <= 0) return;
    int lg = 0;
    while ((1L << (lg + 1)) <= n) ++lg;
    trip_introsort_loop(first, last, 2 * lg);
    if (n > 16) {
        trip_insertion_sort(first, first + 16);
        for (om_trip* i = first + 16; i != last; ++i) trip_linear_insert(i);
    } else {
        trip_insertion_sort(first, last);
    }
}

/* the replayed std::sort over caller-given keys (for tests of the replay
 * against libstdc++ itself): order[] receives the original positions */
void om_trip_sort_keys(const double* keys, long n, int* order)
{
    om_trip* T = (om_trip*)malloc(sizeof(om_trip) * (size_t)(n > 0 ? n : 1));
    for (long i = 0; i < n; ++i) T[i].i = (int)i, T[i].j = T[i].k = 0, T[i].a = keys[i];
    trip_sort(T, T + n);
    for (long i = 0; i < n; ++i) order[i] = T[i].i;
    free(T);
}

int om_best_triple(const double* P, int n, const double* normal, int k, int* ijk)
{
    if (n == 3) {
        ijk[0] = 0, ijk[1] = 1, ijk[2] = 2;
        return 1;
    }
    double M[6];
    double* X = (double*)malloc(sizeof(double) * 2 * (size_t)n);
    projection_matrix(normal, k, M);
    for (int i = 0; i < n; ++i) project(M, P + 3 * i, X + 2 * i);
    const long cap = (long)n * (n - 1) * (n - 2) / 6;
    om_trip* T = (om_trip*)malloc(sizeof(om_trip) * (size_t)cap);
    int m = 0;
    for (int a = 0; a < n; ++a)
        for (int b = a + 1; b < n; ++b)
            for (int c = b + 1; c < n; ++c) {
                double r = om_orient2d(X + 2 * a, X + 2 * b, X + 2 * c);
                if (r == 0.0) continue;
                T[m].i = a, T[m].j = b, T[m].k = c, T[m].a = fabs(r);
                ++m;
            }
    free(X);
    if (m == 0) {
        free(T);
        return 0;
    }
    trip_sort(T, T + m);
    ijk[0] = T[0].i, ijk[1] = T[0].j, ijk[2] = T[0].k;
    free(T);
    return 1;
}

/* compute_segment_plane_intersection_type, math.cpp:391-427 */
char om_segment_type(const double* q, const double* r, const double* P, int n, const int* ijk, int8_t* qs, int8_t* rs)
{
    (void)n;
    const double* a = P + 3 * ijk[0];
    const double* b = P + 3 * ijk[1];
    const double* c = P + 3 * ijk[2];
    int sq = om_orient3d_sign(a, b, c, q);
    int sr = om_orient3d_sign(a, b, c, r);
    if (qs) *qs = (int8_t)sq;
    if (rs) *rs = (int8_t)sr;
    if (sq == 0 && sr == 0) return 'p';
    if (sq == 0) return 'q';
    if (sr == 0) return 'r';
    if (sq == sr) return '0';
    return '1';
}

/* compute_segment_plane_intersection, math.cpp:249-287 (the kernel ignores the
 * return code and keeps p = (0,0,0) when the denominator is zero, kernel.cpp:2505) */
void om_segment_point(const double* normal, double d, const double* q, const double* r, double* p)
{
    double num = d - dot3(q, normal);
    double rq[3] = {r[0] - q[0], r[1] - q[1], r[2] - q[2]};
    double den = dot3(rq, normal);
    if (den == 0.0) return;
    double t = num / den;
    for (int i = 0; i < 3; ++i) p[i] = q[i] + t * (r[i] - q[i]);
}

/* ======================================================================
 * Front end
 * ====================================================================== */

typedef struct {
    uint32_t v0, v1; /* h0 = v0 -> v1 */
    int32_t f0, f1;  /* face of h0 / h1, -1 = none */
} om_edge;

struct om_result {
    double* pbuf;                  /* vertices of one polygon, sized for the largest face */
    uint32_t* ins;                 /* one face's vertex ids, likewise */
    int status, stage;
    int partial;                   /* a successful front end whose registry holds a cut-mesh border edge */
    double com[3], t[3], eps, ms;
    uint32_t Vs, Fs, Vps, Fps;
    double* ps_v;                  /* ps vertex coordinates */
    uint32_t* face_off;            /* ps face -> offset into face_vtx (Fps + 1) */
    uint32_t* face_vtx;            /* ps face vertex lists (targets order) */
    uint32_t* face_edge;           /* ps face edges, parallel to face_vtx */
    om_edge* edges;
    uint32_t ne;
    double* aabb;                  /* per ps face: min xyz, max xyz (cut faces enlarged) */
    uint32_t* pairs;               /* (src, cut) sorted */
    int64_t npairs;
    uint64_t* tests;               /* edge << 32 | face, sorted */
    int64_t ntests;
    int8_t *qs, *rs;
    char *type, *pip;
    uint64_t* reg;                 /* registry keys */
    double* reg_p;
    int64_t nreg;
    uint32_t* cand;                /* candidate faces, sorted */
    int64_t ncand;
    double* plane_n;
    double* plane_d;
    int32_t* plane_k;
    int32_t* plane_ijk;
};

static int cmp_u64(const void* a, const void* b)
{
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y);
}

/* --- edge hash table (open addressing) --- */
typedef struct {
    uint64_t* key;
    uint32_t* val;
    uint64_t mask;
} om_hash;

static void hash_init(om_hash* h, uint64_t n)
{
    uint64_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    h->key = (uint64_t*)malloc(cap * sizeof(uint64_t));
    h->val = (uint32_t*)malloc(cap * sizeof(uint32_t));
    memset(h->key, 0xff, cap * sizeof(uint64_t));
    h->mask = cap - 1;
}

static uint32_t* hash_slot(om_hash* h, uint64_t key, int* found)
{
    uint64_t i = (key * 0x9E3779B97F4A7C15ull) >> 20;
    for (;;) {
        i &= h->mask;
        if (h->key[i] == key) {
            *found = 1;
            return &h->val[i];
        }
        if (h->key[i] == ~0ull) {
            h->key[i] = key;
            *found = 0;
            return &h->val[i];
        }
        ++i;
    }
}

/* hmesh_t::add_face semantics (hmesh.cpp:523-650): edges are created at first
 * use, h0 oriented along that use; a halfedge may carry only one face. */
static int add_face(om_result* R, om_hash* H, uint32_t face, const uint32_t* vi, int n, uint32_t* out_edges)
{
    for (int i = 0; i < n; ++i) {
        uint32_t a = vi[i], b = vi[(i + 1) % n];
        uint64_t key = a < b ? ((uint64_t)a << 32 | b) : ((uint64_t)b << 32 | a);
        int found;
        uint32_t* slot = hash_slot(H, key, &found);
        if (!found) {
            uint32_t e = R->ne++;
            *slot = e;
            R->edges[e].v0 = a;
            R->edges[e].v1 = b;
            R->edges[e].f0 = (int32_t)face;
            R->edges[e].f1 = -1;
            out_edges[i] = e;
        } else {
            om_edge* E = &R->edges[*slot];
            if (E->v0 == a) {
                if (E->f0 != -1) return -1; /* non-manifold */
                E->f0 = (int32_t)face;
            } else {
                if (E->f1 != -1) return -1;
                E->f1 = (int32_t)face;
            }
            out_edges[i] = *slot;
        }
    }
    return 0;
}

/* --- simple BVH over cut-face AABBs, used only to enumerate overlapping pairs
 * (the candidate set does not depend on the tree: bvh.cpp:638-783 returns every
 * (src, cut) face pair whose AABBs overlap under math.h:931-941). --- */
typedef struct {
    double lo[3], hi[3];
    int32_t left, right; /* children, or -1/-(start+2) leaf encoding via count */
    uint32_t start, count;
} om_node;

typedef struct {
    om_node* nodes;
    int nn;
    uint32_t* idx;
    const double* box; /* per item: lo xyz, hi xyz */
} om_bvh;

static const double* g_sort_box;
static int g_sort_axis;
static int cmp_centroid(const void* a, const void* b)
{
    uint32_t i = *(const uint32_t*)a, j = *(const uint32_t*)b;
    double ci = g_sort_box[6 * i + g_sort_axis] + g_sort_box[6 * i + 3 + g_sort_axis];
    double cj = g_sort_box[6 * j + g_sort_axis] + g_sort_box[6 * j + 3 + g_sort_axis];
    return ci < cj ? -1 : (ci > cj);
}

static int bvh_build_rec(om_bvh* B, uint32_t start, uint32_t count)
{
    int id = B->nn++;
    om_node* N = &B->nodes[id];
    for (int a = 0; a < 3; ++a) {
        N->lo[a] = INFINITY;
        N->hi[a] = -INFINITY;
    }
    for (uint32_t i = start; i < start + count; ++i) {
        const double* b = B->box + 6 * B->idx[i];
        for (int a = 0; a < 3; ++a) {
            if (b[a] < N->lo[a]) N->lo[a] = b[a];
            if (b[3 + a] > N->hi[a]) N->hi[a] = b[3 + a];
        }
    }
    N->start = start;
    N->count = count;
    N->left = N->right = -1;
    if (count <= 4) return id;
    int axis = 0;
    double ext = -1;
    for (int a = 0; a < 3; ++a)
        if (N->hi[a] - N->lo[a] > ext) ext = N->hi[a] - N->lo[a], axis = a;
    g_sort_box = B->box;
    g_sort_axis = axis;
    qsort(B->idx + start, count, sizeof(uint32_t), cmp_centroid);
    uint32_t half = count / 2;
    int l = bvh_build_rec(B, start, half);
    int r = bvh_build_rec(B, start + half, count - half);
    B->nodes[id].left = l;
    B->nodes[id].right = r;
    return id;
}

static inline int overlap(const double* a, const double* b)
{
    return (a[0] <= b[3] && a[3] >= b[0]) && (a[1] <= b[4] && a[4] >= b[1]) && (a[2] <= b[5] && a[5] >= b[2]);
}

/* ---------------------------------------------------------------------- */

static uint32_t face_size(const om_mesh* m, uint32_t f) { return m->sz ? m->sz[f] : 3u; }

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

om_result* om_frontend(const om_mesh* src, const om_mesh* cut, const om_options* opt)
{
    double t_start = now_ms();
    om_result* R = (om_result*)calloc(1, sizeof(om_result));
    R->stage = 0;
    const uint32_t Vs = src->nv, Vc = cut->nv, Fs = src->nf, Fc = cut->nf;
    R->Vs = Vs, R->Fs = Fs, R->Vps = Vs + Vc, R->Fps = Fs + Fc;

    /* --- calculate_vertex_parameters, preproc.cpp:2124-2290 --- */
    double bmin[2][3], bmax[2][3], com[2][3];
    const om_mesh* M[2] = {src, cut};
    for (int m = 0; m < 2; ++m) {
        for (int j = 0; j < 3; ++j) {
            bmin[m][j] = 1.7976931348623157e308;
            bmax[m][j] = -1.7976931348623157e308;
            com[m][j] = 0.0;
        }
        for (uint32_t i = 0; i < M[m]->nv; ++i)
            for (int j = 0; j < 3; ++j) {
                double p = M[m]->v[3 * i + j];
                bmax[m][j] = (bmax[m][j] < p) ? p : bmax[m][j]; /* std::max */
                bmin[m][j] = (p < bmin[m][j]) ? p : bmin[m][j]; /* std::min */
                com[m][j] += p;
            }
        for (int j = 0; j < 3; ++j) com[m][j] = com[m][j] / (double)M[m]->nv;
    }
    double gmin[3], tpq[3];
    for (int j = 0; j < 3; ++j) {
        R->com[j] = (com[0][j] + com[1][j]) / 2.0;
        gmin[j] = (bmin[1][j] < bmin[0][j]) ? bmin[1][j] : bmin[0][j];
        tpq[j] = R->com[j] - gmin[j];
    }
    double tl = sqrt(dot3(tpq, tpq));
    for (int j = 0; j < 3; ++j) R->t[j] = tpq[j] + tpq[j] / tl;
    double cd[3];
    for (int j = 0; j < 3; ++j) cd[j] = (bmax[1][j] + R->t[j]) - (bmin[1][j] + R->t[j]);
    double diag = sqrt(dot3(cd, cd));                    /* preproc.cpp:2518 */
    double scalar = opt->eps_absolute ? 1.0 : diag;      /* preproc.cpp:2667-2671 */
    R->eps = scalar * opt->gp_constant;

    /* --- shifted vertices, preproc.cpp:244-276 --- */
    R->ps_v = (double*)malloc(sizeof(double) * 3 * (Vs + Vc));
    double* cut_bvh_v = (double*)malloc(sizeof(double) * 3 * Vc); /* cut coordinates when its BVH was built */
    for (uint32_t i = 0; i < Vs; ++i)
        for (int j = 0; j < 3; ++j) R->ps_v[3 * i + j] = ((src->v[3 * i + j] - R->com[j]) + R->t[j]) + 0.0;
    for (uint32_t i = 0; i < Vc; ++i)
        for (int j = 0; j < 3; ++j) {
            double x = (cut->v[3 * i + j] - R->com[j]) + R->t[j];
            cut_bvh_v[3 * i + j] = x + 0.0;
            R->ps_v[3 * (Vs + i) + j] = x + (opt->has_perturbation ? opt->perturbation[j] : 0.0);
        }

    /* --- polygon soup numbering (kernel.cpp:1593-1732, hmesh.cpp:406-733) --- */
    uint64_t ncorner = 0;
    for (uint32_t f = 0; f < Fs; ++f) ncorner += face_size(src, f);
    for (uint32_t f = 0; f < Fc; ++f) ncorner += face_size(cut, f);
    R->face_off = (uint32_t*)malloc(sizeof(uint32_t) * (R->Fps + 1));
    R->face_vtx = (uint32_t*)malloc(sizeof(uint32_t) * ncorner);
    R->face_edge = (uint32_t*)malloc(sizeof(uint32_t) * ncorner);
    R->edges = (om_edge*)malloc(sizeof(om_edge) * ncorner);
    om_hash H;
    hash_init(&H, ncorner);
    uint32_t off = 0, sidx = 0, maxn = 3;
    for (uint32_t f = 0; f < Fs; ++f) maxn = face_size(src, f) > maxn ? face_size(src, f) : maxn;
    for (uint32_t f = 0; f < Fc; ++f) maxn = face_size(cut, f) > maxn ? face_size(cut, f) : maxn;
    R->ins = (uint32_t*)malloc(sizeof(uint32_t) * maxn); /* one face's vertex list */
    uint32_t* ins = R->ins;
    for (uint32_t f = 0; f < Fs; ++f) {
        uint32_t n = face_size(src, f);
        R->face_off[f] = off;
        for (uint32_t i = 0; i < n; ++i) ins[i] = src->f[sidx + i];
        if (add_face(R, &H, f, ins, (int)n, R->face_edge + off) < 0) {
            R->status = OM_INVALID_SRC_MESH;
            goto done;
        }
        for (uint32_t i = 0; i < n; ++i) R->face_vtx[off + i] = ins[(i + 1) % n];
        sidx += n;
        off += n;
    }
    sidx = 0;
    for (uint32_t f = 0; f < Fc; ++f) {
        uint32_t n = face_size(cut, f), ps_f = Fs + f;
        R->face_off[ps_f] = off;
        /* the cut hmesh lists targets [u1..u0]; ps inserts that list */
        for (uint32_t i = 0; i < n; ++i) ins[i] = Vs + cut->f[sidx + (i + 1) % n];
        if (add_face(R, &H, ps_f, ins, (int)n, R->face_edge + off) < 0) {
            R->status = OM_INVALID_CUT_MESH;
            goto done;
        }
        for (uint32_t i = 0; i < n; ++i) R->face_vtx[off + i] = ins[(i + 1) % n];
        sidx += n;
        off += n;
    }
    R->face_off[R->Fps] = off;

    /* --- face AABBs, bvh.cpp:242-313 (cut faces enlarged by eps) --- */
    R->aabb = (double*)malloc(sizeof(double) * 6 * R->Fps);
    for (uint32_t f = 0; f < R->Fps; ++f) {
        double* b = R->aabb + 6 * f;
        for (int j = 0; j < 3; ++j) {
            b[j] = 1.7976931348623157e308;
            b[3 + j] = -1.7976931348623157e308;
        }
        for (uint32_t c = R->face_off[f]; c < R->face_off[f + 1]; ++c) {
            uint32_t v = R->face_vtx[c];
            const double* p = v < Vs ? R->ps_v + 3 * v : cut_bvh_v + 3 * (v - Vs);
            for (int j = 0; j < 3; ++j) {
                b[3 + j] = (b[3 + j] < p[j]) ? p[j] : b[3 + j];
                b[j] = (p[j] < b[j]) ? p[j] : b[j];
            }
        }
        if (f >= Fs && R->eps > 0.0)
            for (int j = 0; j < 3; ++j) {
                b[3 + j] = b[3 + j] + R->eps;
                b[j] = b[j] - R->eps;
            }
    }

    /* --- candidate pairs: every (src, cut) face pair with overlapping AABBs --- */
    {
        om_bvh B;
        B.nodes = (om_node*)malloc(sizeof(om_node) * (2 * Fc + 8));
        B.nn = 0;
        B.idx = (uint32_t*)malloc(sizeof(uint32_t) * Fc);
        for (uint32_t i = 0; i < Fc; ++i) B.idx[i] = i;
        B.box = R->aabb + 6 * Fs;
        bvh_build_rec(&B, 0, Fc);
        int64_t cap = 1024, n = 0;
        uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * cap);
        int stack[256];
        for (uint32_t s = 0; s < Fs; ++s) {
            const double* sb = R->aabb + 6 * s;
            int sp = 0;
            stack[sp++] = 0;
            while (sp) {
                om_node* N = &B.nodes[stack[--sp]];
                double nb[6] = {N->lo[0], N->lo[1], N->lo[2], N->hi[0], N->hi[1], N->hi[2]};
                if (!overlap(sb, nb)) continue;
                if (N->left < 0) {
                    for (uint32_t i = N->start; i < N->start + N->count; ++i) {
                        uint32_t c = B.idx[i];
                        if (overlap(sb, B.box + 6 * c)) {
                            if (n == cap) keys = (uint64_t*)realloc(keys, sizeof(uint64_t) * (cap *= 2));
                            keys[n++] = (uint64_t)s << 32 | c;
                        }
                    }
                } else {
                    stack[sp++] = N->left;
                    stack[sp++] = N->right;
                }
            }
        }
        qsort(keys, (size_t)n, sizeof(uint64_t), cmp_u64);
        R->npairs = n;
        R->pairs = (uint32_t*)malloc(sizeof(uint32_t) * 2 * (n ? n : 1));
        for (int64_t i = 0; i < n; ++i) {
            R->pairs[2 * i] = (uint32_t)(keys[i] >> 32);
            R->pairs[2 * i + 1] = (uint32_t)keys[i];
        }
        free(keys);
        free(B.nodes);
        free(B.idx);
    }

    /* --- edge-to-face tests (kernel.cpp:1779-1983) and cull (kernel.cpp:2074-2177) --- */
    {
        int64_t cap = 1, n = 0;
        for (int64_t i = 0; i < R->npairs; ++i) {
            uint32_t s = R->pairs[2 * i], c = Fs + R->pairs[2 * i + 1];
            cap += (R->face_off[s + 1] - R->face_off[s]) + (R->face_off[c + 1] - R->face_off[c]);
        }
        uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * cap);
        for (int64_t i = 0; i < R->npairs; ++i) {
            uint32_t s = R->pairs[2 * i], c = Fs + R->pairs[2 * i + 1];
            for (uint32_t k = R->face_off[s]; k < R->face_off[s + 1]; ++k) keys[n++] = (uint64_t)R->face_edge[k] << 32 | c;
            for (uint32_t k = R->face_off[c]; k < R->face_off[c + 1]; ++k) keys[n++] = (uint64_t)R->face_edge[k] << 32 | s;
        }
        qsort(keys, (size_t)n, sizeof(uint64_t), cmp_u64);
        int64_t m = 0;
        for (int64_t i = 0; i < n; ++i) {
            if (i > 0 && keys[i] == keys[i - 1]) continue;
            uint32_t e = (uint32_t)(keys[i] >> 32), g = (uint32_t)keys[i];
            const double* p0 = R->ps_v + 3 * R->edges[e].v1; /* ps.vertex(edge, 0) = target(h0) */
            const double* p1 = R->ps_v + 3 * R->edges[e].v0;
            double eb[6];
            for (int j = 0; j < 3; ++j) {
                double lo = 1.7976931348623157e308, hi = -1.7976931348623157e308;
                hi = (hi < p0[j]) ? p0[j] : hi;
                lo = (p0[j] < lo) ? p0[j] : lo;
                hi = (hi < p1[j]) ? p1[j] : hi;
                lo = (p1[j] < lo) ? p1[j] : lo;
                eb[j] = lo;
                eb[3 + j] = hi;
            }
            if (!overlap(eb, R->aabb + 6 * g)) continue;
            keys[m++] = keys[i];
        }
        R->tests = keys;
        R->ntests = m;
    }

    /* --- face properties for every candidate face, kernel.cpp:2184-2356 --- */
    {
        uint8_t* is_cand = (uint8_t*)calloc(R->Fps, 1);
        for (int64_t i = 0; i < R->npairs; ++i) {
            is_cand[R->pairs[2 * i]] = 1;
            is_cand[Fs + R->pairs[2 * i + 1]] = 1;
        }
        R->ncand = 0;
        for (uint32_t f = 0; f < R->Fps; ++f) R->ncand += is_cand[f];
        R->cand = (uint32_t*)malloc(sizeof(uint32_t) * (R->ncand + 1));
        R->plane_n = (double*)malloc(sizeof(double) * 3 * (R->ncand + 1));
        R->plane_d = (double*)malloc(sizeof(double) * (R->ncand + 1));
        R->plane_k = (int32_t*)malloc(sizeof(int32_t) * (R->ncand + 1));
        R->plane_ijk = (int32_t*)malloc(sizeof(int32_t) * 3 * (R->ncand + 1));
        int64_t c = 0;
        int bad = -1;
        uint32_t maxn = 3;
        for (uint32_t f = 0; f < R->Fps; ++f)
            if (R->face_off[f + 1] - R->face_off[f] > maxn) maxn = R->face_off[f + 1] - R->face_off[f];
        R->pbuf = (double*)malloc(sizeof(double) * 3 * maxn); /* one polygon's vertices */
        double* P = R->pbuf;
        for (uint32_t f = 0; f < R->Fps; ++f) {
            if (!is_cand[f]) continue;
            R->cand[c] = f;
            uint32_t n = R->face_off[f + 1] - R->face_off[f];
            for (uint32_t i = 0; i < n; ++i)
                memcpy(P + 3 * i, R->ps_v + 3 * R->face_vtx[R->face_off[f] + i], 3 * sizeof(double));
            R->plane_k[c] = om_plane(P, (int)n, R->plane_n + 3 * c, R->plane_d + c);
            double* nn = R->plane_n + 3 * c;
            if (dot3(nn, nn) == 0.0 || isnan(nn[0]) || isnan(nn[1]) || isnan(nn[2])) bad = (int)f;
            R->plane_ijk[3 * c] = 0, R->plane_ijk[3 * c + 1] = 1, R->plane_ijk[3 * c + 2] = 2;
            if (bad < 0 && n > 3) {
                int rc = om_best_triple(P, (int)n, nn, R->plane_k[c], R->plane_ijk + 3 * c);
                if (rc < 0) R->status = OM_UNSUPPORTED;
                if (rc == 0) R->plane_ijk[3 * c] = -1; /* all collinear: type '0' */
            }
            ++c;
        }
        free(is_cand);
        if (bad >= 0) {
            /* kernel.cpp:2320-2331: "tmp_local > sm_face_count" picks the mesh */
            R->status = ((uint32_t)bad > Fs) ? OM_INVALID_CUT_MESH : OM_INVALID_SRC_MESH;
            goto done;
        }
        if (R->status == OM_UNSUPPORTED) goto done;
    }

    /* --- narrow phase, kernel.cpp:2456-2658 --- */
    R->stage = 1;
    {
        int64_t T = R->ntests;
        R->qs = (int8_t*)malloc(T + 1);
        R->rs = (int8_t*)malloc(T + 1);
        R->type = (char*)malloc(T + 1);
        R->pip = (char*)malloc(T + 1);
        R->reg = (uint64_t*)malloc(sizeof(uint64_t) * (T + 1));
        R->reg_p = (double*)malloc(sizeof(double) * 3 * (T + 1));
        R->nreg = 0;
        int gp = 0;
        for (int64_t i = 0; i < T; ++i) {
            uint32_t e = (uint32_t)(R->tests[i] >> 32), g = (uint32_t)R->tests[i];
            /* candidate face g -> its plane slot */
            int64_t lo = 0, hi = R->ncand - 1;
            while (lo < hi) {
                int64_t mid = (lo + hi) / 2;
                if (R->cand[mid] < g) lo = mid + 1;
                else hi = mid;
            }
            const double* nrm = R->plane_n + 3 * lo;
            double d = R->plane_d[lo];
            int k = R->plane_k[lo];
            const int* ijk = R->plane_ijk + 3 * lo;
            uint32_t n = R->face_off[g + 1] - R->face_off[g];
            double* P = R->pbuf;
            for (uint32_t j = 0; j < n; ++j)
                memcpy(P + 3 * j, R->ps_v + 3 * R->face_vtx[R->face_off[g] + j], 3 * sizeof(double));
            const double* q = R->ps_v + 3 * R->edges[e].v0; /* source(h0) */
            const double* r = R->ps_v + 3 * R->edges[e].v1; /* target(h0) */
            char ty;
            if (ijk[0] < 0) {
                ty = '0';
                R->qs[i] = R->rs[i] = 0;
            } else {
                ty = om_segment_type(q, r, P, (int)n, ijk, &R->qs[i], &R->rs[i]);
            }
            R->type[i] = ty;
            R->pip[i] = '-';
            if (ty == '0') continue;
            if (ty != '1') {
                const double* touch[2];
                int nt = 0;
                if (ty == 'q') touch[nt++] = q;
                else if (ty == 'r') touch[nt++] = r;
                else {
                    touch[nt++] = q;
                    touch[nt++] = r;
                }
                for (int j = 0; j < nt; ++j) {
                    char res = om_point_in_polygon(touch[j], P, (int)n, nrm, k);
                    R->pip[i] = res;
                    if (res == 'i' || res == 'v' || res == 'e') {
                        gp = 1;
                        break;
                    }
                }
                continue;
            }
            double p[3] = {0.0, 0.0, 0.0};
            om_segment_point(nrm, d, q, r, p);
            char res = om_point_in_polygon(p, P, (int)n, nrm, k);
            R->pip[i] = res;
            if (res == 'v' || res == 'e') {
                gp = 1;
                continue;
            }
            if (res == 'i') {
                R->reg[R->nreg] = R->tests[i];
                memcpy(R->reg_p + 3 * R->nreg, p, 3 * sizeof(double));
                R->nreg++;
            }
        }
        if (gp) {
            R->status = OM_GENERAL_POSITION_VIOLATION;
            goto done;
        }
    }

    /* --- empty registry, kernel.cpp:3211-3229 --- */
    R->stage = 2;
    if (R->nreg == 0) {
        R->status = (opt->enforce_gp && opt->enforcement_count > 0) ? OM_GENERAL_POSITION_VIOLATION : OM_SUCCESS;
        goto done;
    }
    /* --- partial cuts with through-cuts required, kernel.cpp:3269-3278 --- */
    R->stage = 3;
    {
        int partial = 0;
        for (int64_t i = 0; i < R->nreg && !partial; ++i) {
            const om_edge* E = &R->edges[R->reg[i] >> 32];
            partial = E->v0 >= Vs && (E->f0 < 0 || E->f1 < 0);
        }
        if (partial && opt->require_through_cuts) {
            R->status = OM_SUCCESS;
            R->partial = 1;
            goto done;
        }
        R->partial = partial; /* cleared again below if the pair check fails */
    }
    /* --- every (src face, cut face) pair needs >= 2 points, kernel.cpp:3439-3452 --- */
    {
        uint64_t* k = (uint64_t*)malloc(sizeof(uint64_t) * 2 * R->nreg);
        int64_t n = 0;
        for (int64_t i = 0; i < R->nreg; ++i) {
            const om_edge* E = &R->edges[R->reg[i] >> 32];
            uint32_t g = (uint32_t)R->reg[i];
            int on_cut = E->f0 >= (int32_t)Fs; /* tested_edge_face = face(h0) */
            if (on_cut) {
                k[n++] = (uint64_t)g << 32 | (uint32_t)E->f0;
                if (E->f1 >= 0) k[n++] = (uint64_t)g << 32 | (uint32_t)E->f1;
            } else {
                k[n++] = (uint64_t)(uint32_t)E->f0 << 32 | g;
                if (E->f1 >= 0) k[n++] = (uint64_t)(uint32_t)E->f1 << 32 | g;
            }
        }
        qsort(k, (size_t)n, sizeof(uint64_t), cmp_u64);
        int bad = 0;
        for (int64_t i = 0; i < n;) {
            int64_t j = i;
            while (j < n && k[j] == k[i]) ++j;
            if (j - i < 2) bad = 1;
            i = j;
        }
        free(k);
        if (bad) {
            R->status = opt->enforce_gp ? OM_GENERAL_POSITION_VIOLATION : OM_INVALID_MESH_INTERSECTION;
            R->partial = 0;
            goto done;
        }
    }
    R->stage = 4;
    R->status = OM_SUCCESS;
done:
    free(H.key);
    free(H.val);
    free(cut_bvh_v);
    R->ms = now_ms() - t_start;
    return R;
}

/* ---------------------------------------------------------------------- */

void om_free(om_result* R)
{
    if (!R) return;
    void* ptrs[] = {R->ps_v, R->face_off, R->face_vtx, R->face_edge, R->edges, R->aabb, R->pairs, R->tests,
                    R->qs, R->rs, R->type, R->pip, R->reg, R->reg_p, R->cand, R->plane_n, R->plane_d,
                    R->plane_k, R->plane_ijk, R->pbuf, R->ins};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); ++i) free(ptrs[i]);
    free(R);
}

int om_status(const om_result* R) { return R->status; }
int om_stage(const om_result* R) { return R->stage; }
int om_partial(const om_result* R) { return R->partial; }
double om_time_ms(const om_result* R) { return R->ms; }

void om_shift(const om_result* R, double com[3], double t[3], double* eps)
{
    memcpy(com, R->com, sizeof(R->com));
    memcpy(t, R->t, sizeof(R->t));
    *eps = R->eps;
}

int64_t om_count(const om_result* R, int what)
{
    switch (what) {
    case OM_N_PAIRS: return R->npairs;
    case OM_N_TESTS: return R->qs ? R->ntests : 0;
    case OM_N_REGISTRY: return R->nreg;
    case OM_N_EDGES: return R->ne;
    case OM_N_CAND_FACES: return R->ncand;
    }
    return -1;
}

void om_copy_pairs(const om_result* R, uint32_t* pairs)
{
    memcpy(pairs, R->pairs, sizeof(uint32_t) * 2 * R->npairs);
}

void om_copy_tests(const om_result* R, uint32_t* edge, uint32_t* face, int8_t* qs, int8_t* rs, char* type, char* pip)
{
    for (int64_t i = 0; i < R->ntests; ++i) {
        edge[i] = (uint32_t)(R->tests[i] >> 32);
        face[i] = (uint32_t)R->tests[i];
    }
    if (R->qs) {
        memcpy(qs, R->qs, R->ntests);
        memcpy(rs, R->rs, R->ntests);
        memcpy(type, R->type, R->ntests);
        memcpy(pip, R->pip, R->ntests);
    }
}

void om_copy_registry(const om_result* R, uint32_t* edge, uint32_t* face, double* points)
{
    for (int64_t i = 0; i < R->nreg; ++i) {
        edge[i] = (uint32_t)(R->reg[i] >> 32);
        face[i] = (uint32_t)R->reg[i];
    }
    memcpy(points, R->reg_p, sizeof(double) * 3 * R->nreg);
}

void om_copy_edges(const om_result* R, uint32_t* v0, uint32_t* v1, int32_t* f0, int32_t* f1)
{
    for (uint32_t e = 0; e < R->ne; ++e) {
        v0[e] = R->edges[e].v0;
        v1[e] = R->edges[e].v1;
        f0[e] = R->edges[e].f0;
        f1[e] = R->edges[e].f1;
    }
}

void om_copy_planes(const om_result* R, uint32_t* face, double* normal, double* d, int32_t* k, int32_t* triple)
{
    memcpy(face, R->cand, sizeof(uint32_t) * R->ncand);
    memcpy(normal, R->plane_n, sizeof(double) * 3 * R->ncand);
    memcpy(d, R->plane_d, sizeof(double) * R->ncand);
    memcpy(k, R->plane_k, sizeof(int32_t) * R->ncand);
    memcpy(triple, R->plane_ijk, sizeof(int32_t) * 3 * R->ncand);
}

void om_copy_ps_vertices(const om_result* R, double* xyz)
{
    memcpy(xyz, R->ps_v, sizeof(double) * 3 * R->Vps);
}

/* ======================================================================
 * Cut-path edges, kernel.cpp:3427-3615 (grouping: kernel.cpp:2616-2631;
 * linear_projection_sort: kernel.cpp:1496-1531; registry faces:
 * ps_get_ivtx_registry_entry_faces, kernel.cpp:1407-1415)
 * ====================================================================== */

typedef struct {
    uint32_t sm, cm, k;
} om_item;

static int cmp_item(const void* a, const void* b)
{
    const om_item* x = (const om_item*)a;
    const om_item* y = (const om_item*)b;
    if (x->sm != y->sm) return x->sm < y->sm ? -1 : 1;
    if (x->cm != y->cm) return x->cm < y->cm ? -1 : 1;
    return x->k < y->k ? -1 : (x->k > y->k);
}

static int64_t plane_slot(const om_result* R, uint32_t g)
{
    int64_t lo = 0, hi = R->ncand - 1;
    while (lo < hi) {
        int64_t mid = (lo + hi) / 2;
        if (R->cand[mid] < g) lo = mid + 1;
        else hi = mid;
    }
    return (R->ncand > 0 && R->cand[lo] == g) ? lo : -1;
}

/* The cut-path edges of R's registry taken in `order` (order[k] = registry
 * index of point k; NULL: registry order).  Writes up to `cap` edges (source
 * and target ranks, the two faces each is filed under, its group and kind)
 * and returns their count; *short_group = 1 when a group holds one point
 * (the dispatch stops there and no edge is meaningful), -1 on a midpoint
 * test against a face without a plane. */
int64_t om_cutpath(const om_result* R, const uint32_t* order, uint32_t* src, uint32_t* tgt, uint32_t* faces,
                   uint32_t* groups, uint8_t* kinds, int64_t cap, int* short_group)
{
    const int64_t n = R->nreg;
    *short_group = 0;
    om_item* it = (om_item*)malloc(sizeof(om_item) * (size_t)(2 * n + 1));
    int64_t ni = 0;
    for (int64_t k = 0; k < n; ++k) {
        const uint64_t i = order ? order[k] : (uint64_t)k;
        const uint32_t e = (uint32_t)(R->reg[i] >> 32), g = (uint32_t)R->reg[i];
        const int32_t f0 = R->edges[e].f0, f1 = R->edges[e].f1;
        const int32_t ef = f0 >= 0 ? f0 : f1;
        if ((uint32_t)ef >= R->Fs) { /* a cut-mesh edge through source face g */
            it[ni++] = (om_item){g, (uint32_t)ef, (uint32_t)k};
            const int32_t other = ef == f0 ? f1 : -1;
            if (other >= 0) it[ni++] = (om_item){g, (uint32_t)other, (uint32_t)k};
        } else { /* a source-mesh edge through cut face g */
            it[ni++] = (om_item){(uint32_t)ef, g, (uint32_t)k};
            const int32_t other = ef == f0 ? f1 : f0;
            if (other >= 0) it[ni++] = (om_item){(uint32_t)other, g, (uint32_t)k};
        }
    }
    qsort(it, (size_t)ni, sizeof(om_item), cmp_item); /* std::map order; points in registry order */
    /* made edges: a hash on the (min, max) rank pair */
    om_hash made;
    hash_init(&made, (uint64_t)ni + 1);
    int64_t ne = 0;
    uint32_t group = 0;
    double* proj = (double*)malloc(sizeof(double) * (size_t)(ni + 1));
    int* ord = (int*)malloc(sizeof(int) * (size_t)(ni + 1));
    uint32_t* pts = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(ni + 1));
    for (int64_t a = 0; a < ni; ++group) {
        int64_t b = a;
        while (b < ni && it[b].sm == it[a].sm && it[b].cm == it[a].cm) ++b;
        const int64_t len = b - a;
        const uint32_t sm = it[a].sm, cm = it[a].cm;
        if (len < 2) {
            *short_group = 1;
            break;
        }
        for (int64_t q = 0; q < len; ++q) pts[q] = it[a + q].k;
        if (len > 2) { /* linear_projection_sort */
            const uint32_t i0 = order ? order[pts[0]] : pts[0];
            const uint32_t i1 = order ? order[pts[1]] : pts[1];
            const double* o = R->reg_p + 3 * (size_t)i0;
            const double* d = R->reg_p + 3 * (size_t)i1;
            double v[3], u[3], sq = 0.0;
            for (int j = 0; j < 3; ++j) v[j] = o[j] - d[j];
            for (int j = 0; j < 3; ++j) sq = sq + v[j] * v[j];
            const double lv = sqrt(sq);
            for (int j = 0; j < 3; ++j) u[j] = v[j] / lv;
            for (int64_t q = 0; q < len; ++q) {
                const uint32_t iq = order ? order[pts[q]] : pts[q];
                double w[3], pr = 0.0;
                for (int j = 0; j < 3; ++j) w[j] = o[j] - R->reg_p[3 * (size_t)iq + j];
                for (int j = 0; j < 3; ++j) pr = pr + w[j] * u[j];
                proj[q] = -pr; /* the replayed std::sort orders descending keys: negated, ascending */
            }
            om_trip_sort_keys(proj, (long)len, ord);
            uint32_t* tmp = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)len);
            for (int64_t q = 0; q < len; ++q) tmp[q] = pts[ord[q]];
            memcpy(pts, tmp, sizeof(uint32_t) * (size_t)len);
            free(tmp);
        }
        for (int64_t q = 1; q < len; ++q) {
            const uint32_t s_ = pts[q - 1], t_ = pts[q];
            const uint64_t key = s_ < t_ ? ((uint64_t)s_ << 32 | t_) : ((uint64_t)t_ << 32 | s_);
            int found = 0;
            uint32_t* slot = hash_slot(&made, key, &found);
            if (!found) *slot = 0;     /* a pair seen for the first time: not made yet */
            if (*slot) continue;       /* m0.halfedge(src, tgt) exists */
            uint32_t fa = sm, fb = cm;
            if (len > 2) {
                const uint32_t ia = order ? order[s_] : s_, ib = order ? order[t_] : t_;
                int32_t FA[3], FB[3];
                FA[0] = (int32_t)(uint32_t)R->reg[ia], FA[1] = R->edges[R->reg[ia] >> 32].f0, FA[2] = R->edges[R->reg[ia] >> 32].f1;
                FB[0] = (int32_t)(uint32_t)R->reg[ib], FB[1] = R->edges[R->reg[ib] >> 32].f0, FB[2] = R->edges[R->reg[ib] >> 32].f1;
                double mid[3];
                for (int j = 0; j < 3; ++j) mid[j] = (R->reg_p[3 * (size_t)ib + j] + R->reg_p[3 * (size_t)ia + j]) * 0.5;
                int inside = 0;
                uint32_t in_face[3] = {0, 0, 0};
                for (int x = 0; x < 3; ++x) {
                    const int32_t f = FA[x];
                    if (f < 0 || !(f == FB[0] || f == FB[1] || f == FB[2])) continue;
                    const int64_t ps = plane_slot(R, (uint32_t)f);
                    if (ps < 0) {
                        *short_group = -1;
                        continue;
                    }
                    const uint32_t nv = R->face_off[f + 1] - R->face_off[f];
                    double* P = (double*)malloc(sizeof(double) * 3 * nv);
                    for (uint32_t j = 0; j < nv; ++j)
                        memcpy(P + 3 * j, R->ps_v + 3 * (size_t)R->face_vtx[R->face_off[f] + j], 3 * sizeof(double));
                    const char res = om_point_in_polygon(mid, P, (int)nv, R->plane_n + 3 * ps, R->plane_k[ps]);
                    free(P);
                    if (res == 'i') {
                        if (inside < 3) in_face[inside] = (uint32_t)f;
                        ++inside;
                    }
                }
                if (inside != 2) continue;
                fa = in_face[0];
                fb = in_face[1];
            }
            *slot = 1; /* the edge is made */
            if (ne < cap) {
                if (src) src[ne] = s_;
                if (tgt) tgt[ne] = t_;
                if (faces) faces[2 * ne] = fa, faces[2 * ne + 1] = fb;
                if (groups) groups[ne] = group;
                if (kinds) kinds[ne] = len > 2 ? 1 : 0;
            }
            ++ne;
        }
        a = b;
    }
    free(made.key);
    free(made.val);
    free(proj);
    free(ord);
    free(pts);
    free(it);
    return ne;
}
