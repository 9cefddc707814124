/*
 * mcrt_oracle.c -- CPU FP64 restatement of the reference photon-packet path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Each function cites the reference routine it restates (paths relative to
 * the reference's src3_mpi_cfitsio/ directory).  Operation order follows the
 * Fortran so that FP64 results are reproducible bit for bit by an independent
 * implementation with floating-point contraction disabled.
 *
 * Deliberate, documented deviations from the reference (all outside the
 * arithmetic of the path):
 *  - ran() is either the reference's L'Ecuyer stream or a counter-based
 *    Philox4x32-10 stream per packet (the GPU's RNG);
 *  - a packet re-emitted by the stellar surface samples its frequency from
 *    the blackbody (bbfreq) instead of a tabulated atmosphere (atmosfreq),
 *    because the synthetic star is a blackbody;
 *  - fatal `stop`s become per-packet error counters;
 *  - findangle3 leaves ri1 unset at the poles; here it is 0.
 */
#include "mcrt_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* constants.f90 */
static const double PI = 3.1415926535897932384626433;
static const double R2P = 2.0 * 3.1415926535897932384626433;
static const double RSOL = 6.9598e10;
static const double TMAXK = 199999.0;

#define CELL(m, ir, it, ip) ((size_t)((ir) - 1) + (size_t)((m)->nrg - 1) * ((size_t)((it) - 1) + (size_t)((m)->ntg - 1) * (size_t)((ip) - 1)))

/* ======================================================================
 * Random numbers
 * ====================================================================== */

/* Philox4x32-10 (Salmon et al., SC'11, "Parallel random numbers: as easy as
 * 1, 2, 3"): 10 rounds of the 4x32 multiply-xor network with Weyl key bumps. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Stream layout shared with the GPU: key = seed, counter = (n/4, packet lo,
 * packet hi, 'MCRT'), word n%4. */
uint32_t orc_packet_word(uint64_t seed, uint64_t packet, uint32_t n)
{
    uint32_t ctr[4] = { n >> 2, (uint32_t)packet, (uint32_t)(packet >> 32), 0x4D435254u };
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    uint32_t out[4];
    orc_philox4x32_10(ctr, key, out);
    return out[n & 3u];
}

/* L'Ecuyer & Cote (ACM TOMS 17:98-111, 1991) as vendored in
 * ecuyer_cote_mod.f90: a1=40014, m1=2147483563, a2=40692, m2=2147483399,
 * 32 generators split by a^(2^(v+w)) (a1vw, a2vw). */
#define LEC_NG 32
static int32_t lec_cg1[LEC_NG], lec_cg2[LEC_NG];
static int lec_cur = 0;

static int32_t lec_mulmod(int32_t a, int32_t s, int32_t m)
{   /* exact (a*s) mod m; multiply_modulo (ecuyer_cote_mod.f90) */
    return (int32_t)(((int64_t)a * (int64_t)s) % (int64_t)m);
}

void orc_lecuyer_set_all_seeds(int32_t s1, int32_t s2)
{   /* set_all_seeds + reinitialize_current_generator(-1) for every g */
    lec_cg1[0] = s1; lec_cg2[0] = s2;
    for (int g = 1; g < LEC_NG; ++g) {
        lec_cg1[g] = lec_mulmod(2082007225, lec_cg1[g - 1], 2147483563);
        lec_cg2[g] = lec_mulmod(784306273, lec_cg2[g - 1], 2147483399);
    }
    lec_cur = 0;
}

void orc_lecuyer_set_generator(int32_t g) { lec_cur = g - 1; }

double orc_lecuyer_ran(void)
{   /* random_large_integer (Schrage steps) * 4.656613057d-10 */
    int32_t s1 = lec_cg1[lec_cur], s2 = lec_cg2[lec_cur], k, z;
    k = s1 / 53668;
    s1 = 40014 * (s1 - k * 53668) - k * 12211;
    if (s1 < 0) s1 += 2147483563;
    k = s2 / 52774;
    s2 = 40692 * (s2 - k * 52774) - k * 3791;
    if (s2 < 0) s2 += 2147483399;
    lec_cg1[lec_cur] = s1; lec_cg2[lec_cur] = s2;
    z = s1 - s2;
    if (z < 1) z += 2147483563 - 1;
    return (double)z * 4.656613057e-10;
}

typedef struct {
    int mode;              /* 0 philox, 1 lecuyer */
    uint64_t seed, packet;
    uint32_t n;
} orc_rng;

static double ran(orc_rng *r)
{
    if (r->mode == 1) return orc_lecuyer_ran();
    uint32_t w = orc_packet_word(r->seed, r->packet, r->n++);
    return ((double)w + 0.5) * 2.3283064365386963e-10;
}

/* ======================================================================
 * closest_wall.f90:13-66
 * ====================================================================== */
void orc_reset_t(orc_walls *w)
{
    w->tn1 = -1.7976931348623157e308; w->tp1 = 1.7976931348623157e308; w->tp2 = 1.7976931348623157e308;
    w->in1 = 0; w->ip1 = 0; w->ip2 = 0;
}

void orc_insert_t(orc_walls *w, double t, int32_t i)
{
    if (t < 0.0) {
        if (t > w->tn1) { w->tn1 = t; w->in1 = i; }
    } else {
        if (t < w->tp1) { w->tp2 = w->tp1; w->ip2 = w->ip1; w->tp1 = t; w->ip1 = i; }
        else if (t < w->tp2) { w->tp2 = t; w->ip2 = i; }
    }
}

int orc_find_next_wall(const orc_walls *w, int on_wall, double *t, int32_t *i)
{
    if (on_wall) {
        if (fabs(w->tn1) < fabs(w->tp1)) { *t = w->tp1; *i = w->ip1; }
        else { *t = w->tp2; *i = w->ip2; }
    } else { *t = w->tp1; *i = w->ip1; }
    return (*t == 1.7976931348623157e308) ? -1 : 0;   /* "ERROR - no valid solution" */
}

/* radface.f90:44-62 */
void orc_radface(orc_walls *w, double ux, double uy, double uz, double x1, double y1, double z1, double R2_1, double R2_2)
{
    double bb = 2.0 * (x1 * ux + y1 * uy + uz * z1);
    double cc = x1 * x1 + y1 * y1 + z1 * z1 - R2_1;
    double descr = bb * bb - 4.0 * cc;
    if (descr > 0.0) {
        descr = sqrt(descr);
        orc_insert_t(w, (-bb + descr) * 0.5, 1);
        orc_insert_t(w, (-bb - descr) * 0.5, 1);
    }
    cc = x1 * x1 + y1 * y1 + z1 * z1 - R2_2;
    descr = bb * bb - 4.0 * cc;
    if (descr > 0.0) {
        descr = sqrt(descr);
        orc_insert_t(w, (-bb + descr) * 0.5, 2);
        orc_insert_t(w, (-bb - descr) * 0.5, 2);
    }
}

/* thetaface.f90:59-115 */
void orc_thetaface(orc_walls *w, double ux, double uy, double uz, double x1, double y1, double z1, double t1, double t2)
{
    double tmp1 = x1 * ux + y1 * uy, tmp2 = ux * ux + uy * uy, tmp3 = x1 * x1 + y1 * y1;
    double ta, tb, aa, bb, cc, descr;
    int32_t i1, i2;
    if (t2 < t1) { ta = t2; tb = t1; i1 = 4; i2 = 3; }
    else { ta = t1; tb = t2; i1 = 3; i2 = 4; }
    if (ta < 0.0) {
        if (uz != 0.0) orc_insert_t(w, -z1 / uz, i1);
    } else {
        bb = 2.0 * (tmp1 - ta * uz * z1);
        aa = tmp2 - ta * uz * uz;
        cc = tmp3 - z1 * z1 * ta;
        if (fabs(aa) > 1.0e-10) {
            descr = bb * bb - 4.0 * aa * cc;
            if (descr > 0.0 && ta != 0.0) {
                descr = sqrt(descr);
                orc_insert_t(w, (-bb + descr) * 0.5 / aa, i1);
                orc_insert_t(w, (-bb - descr) * 0.5 / aa, i1);
            }
        } else {
            if (fabs(bb) > 1.0e-10) orc_insert_t(w, -cc / bb, i1);
        }
    }
    bb = 2.0 * (tmp1 - tb * uz * z1);
    aa = tmp2 - tb * uz * uz;
    cc = tmp3 - z1 * z1 * tb;
    if (fabs(aa) > 1.0e-10) {
        descr = bb * bb - 4.0 * aa * cc;
        if (descr > 0.0 && tb != 0.0) {
            descr = sqrt(descr);
            orc_insert_t(w, (-bb + descr) * 0.5 / aa, i2);
            orc_insert_t(w, (-bb - descr) * 0.5 / aa, i2);
        }
    } else {
        if (fabs(bb) > 1.0e-10) orc_insert_t(w, -cc / bb, i2);
    }
}

/* phiface.f90:48-56 with C=D=0 (tauint_3d_2_fuv.f90:47) */
void orc_phiface(orc_walls *w, double A1, double B1, double A2, double B2, double ux, double uy, double uz, double x1, double y1, double z1)
{
    double denom1 = (A1 * ux + B1 * uy + 0.0 * uz);
    if (denom1 != 0.0) orc_insert_t(w, -(A1 * x1 + B1 * y1 + 0.0 * z1 + 0.0) / denom1, 5);
    double denom2 = (A2 * ux + B2 * uy + 0.0 * uz);
    if (denom2 != 0.0) orc_insert_t(w, -(A2 * x1 + B2 * y1 + 0.0 * z1 + 0.0) / denom2, 6);
}

/* locate, nr_2.f:131-154 (1-based result; 0 or n when outside) */
int32_t orc_locate(const double *xx, int32_t n, double x)
{
    int32_t jl = 0, ju = n + 1, jm;
    int ascnd = xx[n - 1] > xx[0];
    while (ju - jl > 1) {
        jm = (ju + jl) / 2;
        if (ascnd == (x > xx[jm - 1])) jl = jm; else ju = jm;
    }
    if (jl == 0 && x == xx[0]) jl = 1;
    return jl;
}

/* check.f90:1-30 */
static double check(double x) { if (x > 1.0) return 1.0; if (x < -1.0) return -1.0; return x; }

/* findangle3.f90:31-67 */
void orc_findangle3(double cost1, double sint1, double phi1, double cosp1, double sinp1,
                    double cost2, double sint2, double phi2, double cosp2, double sinp2,
                    double *costnew, double *sintnew, double *phinew, double *ri1)
{
    double cn = cost1 * cost2 + (sint1 * sint2 * (cosp1 * cosp2 + sinp1 * sinp2));
    cn = check(cn);
    double sn = sqrt(1.0 - cn * cn);
    *costnew = cn; *sintnew = sn; *ri1 = 0.0;
    if (sint1 > 1.0e-8) {
        if (sn > 1.0e-8) {
            double cosx = (cost2 - (cost1 * cn)) / (sint1 * sn);
            cosx = check(cosx);
            double x = acos(cosx), diff;
            if (phi2 > phi1) {
                diff = phi2 - phi1;
                if (diff > PI) { *ri1 = x; *phinew = PI + x; }
                else { *phinew = PI - x; *ri1 = R2P - x; }
            } else {
                diff = phi1 - phi2;
                if (diff > PI) { *phinew = PI - x; *ri1 = R2P - x; }
                else { *phinew = PI + x; *ri1 = x; }
            }
            (void)diff;
        } else {
            *phinew = 0.0;
        }
    } else {
        *phinew = phi2;
    }
}

/* dustmat.f90:42-67 (modified Cornette-Shanks p1, White 1979 p2..p4) */
void orc_dustmat(double g, double pl, double cost, double cost2, double p[4])
{
    double g2 = g * g, g2p1 = g2 + 1.0, onemg2 = 1.0 - g2, twog = 2.0 * g;
    double temp = g2p1 - twog * cost;
    double p1 = 1.5 * onemg2 / (2.0 + g2) * (1.0 + cost2) / (temp * sqrt(temp));
    p[0] = p1;
    p[1] = -pl * p1 * (1.0 - cost2) / (1.0 + cost2);
    p[2] = p1 * 2.0 * cost / (1.0 + cost2);
    p[3] = 0.0;
}

static double cs_peak(double g)
{   /* stokes.f90:54-70 */
    double g2 = g * g, g2p1 = g2 + 1.0, onemg2 = 1.0 - g2, twog = 2.0 * g, temp, peak;
    if (g > 0.0) { temp = g2p1 - twog; peak = 1.5 * (onemg2) / (2.0 + g2) * 2.0 / temp / sqrt(temp); }
    else { temp = g2p1 + twog; peak = 1.5 * (onemg2) / (2.0 + g2) * 2.0 / temp / sqrt(temp); }
    return peak * 1.01;
}

/* Mueller rotation of stokes.f90:117-336 for an accepted (bmu, ri1).
 * dir = {cost, sint, phi, cosp, sinp}. */
void orc_stokes_apply(double g, double pl, double bmu, double ri1, double dir[5], double s[4])
{
    double costp = dir[0], sintp = dir[1], phip = dir[2];
    double cost = dir[0], sint = dir[1], phi = dir[2];
    double sip = s[0], sqp = s[1], sup = s[2], svp = s[3];
    double p[4], cosb2 = bmu * bmu, sinbt;
    double a11, a12, a13, a21, a22, a23, a24, a31, a32, a33, a34, a42, a43, a44, a;
    double sini2 = 0.0, cosi2 = 0.0, bott, cosdph, sin2i2, cos2i2, sin2, cos2, sin2cos1, cos2sin1;
    orc_dustmat(g, pl, bmu, cosb2, p);
    sinbt = sqrt(1.0 - cosb2);
    if (ri1 > PI) {
        double ri3 = R2P - ri1, cosi3 = cos(ri3), sini3 = sin(ri3);
        double sin2i3 = 2.0 * sini3 * cosi3, cos2i3 = 2.0 * (cosi3 * cosi3) - 1.0;
        a11 = p[0]; a12 = p[1] * cos2i3; a13 = p[1] * sin2i3;
        a = (a11 * sip + a12 * sqp + a13 * sup) / sip;
        if (bmu == 1.0) return;
        else if (bmu == -1.0) { sini3 = 1.0; sini2 = 1.0; cosi3 = 0.0; cosi2 = 0.0; cost = -costp; sint = sintp; }
        else {
            cost = costp * bmu + sintp * sinbt * cosi3;
            if (fabs(cost) < 1.0) {
                sint = fabs(sqrt(1.0 - cost * cost));
                sini2 = sini3 * sintp / sint;
                bott = sint * sinbt;
                cosi2 = costp / bott - cost * bmu / bott;
            } else {
                sint = 0.0; sini2 = 0.0;
                if (cost >= 1.0) cosi2 = -1.0;
                if (cost <= -1.0) cosi2 = 1.0;
            }
        }
        cosdph = -cosi2 * cosi3 + sini2 * sini3 * bmu;
        if (fabs(cosdph) > 1.0) cosdph = (cosdph > 1.0) ? 1.0 : -1.0;
        phi = phip + acos(cosdph);
        if (phi > R2P) phi = phi - R2P;
        if (phi < 0.0) phi = phi + R2P;
        sin2i2 = 2.0 * sini2 * cosi2; cos2i2 = 2.0 * (cosi2 * cosi2) - 1.0;
        sin2 = sin2i2 * sin2i3; cos2 = cos2i2 * cos2i3;
        sin2cos1 = sin2i2 * cos2i3; cos2sin1 = cos2i2 * sin2i3;
        a21 = p[1] * cos2i2; a22 = p[0] * cos2 - p[2] * sin2; a23 = p[0] * cos2sin1 + p[2] * sin2cos1; a24 = -p[3] * sin2i2;
        a31 = -p[1] * sin2i2; a32 = -p[0] * sin2cos1 - p[2] * cos2sin1; a33 = -p[0] * sin2 + p[2] * cos2; a34 = -p[3] * cos2i2;
        a42 = -p[3] * sin2i3; a43 = p[3] * cos2i3; a44 = p[2];
    } else {
        double cosi1 = cos(ri1), sini1 = sin(ri1);
        double sin2i1 = 2.0 * sini1 * cosi1, cos2i1 = 2.0 * (cosi1 * cosi1) - 1.0;
        a11 = p[0]; a12 = p[1] * cos2i1; a13 = -p[1] * sin2i1;
        a = (a11 * sip + a12 * sqp + a13 * sup) / sip;
        if (bmu == 1.0) return;
        else if (bmu == -1.0) { sini1 = 1.0; sini2 = 1.0; cosi1 = 0.0; cosi2 = 0.0; cost = -costp; sint = sintp; }
        else {
            cost = costp * bmu + sintp * sinbt * cosi1;
            if (fabs(cost) < 1.0) {
                sint = fabs(sqrt(1.0 - cost * cost));
                sini2 = sini1 * sintp / sint;
                bott = sint * sinbt;
                cosi2 = costp / bott - cost * bmu / bott;
            } else {
                sint = 0.0; sini2 = 0.0;
                if (cost >= 1.0) cosi2 = -1.0;
                if (cost <= -1.0) cosi2 = 1.0;
            }
        }
        cosdph = -cosi1 * cosi2 + sini1 * sini2 * bmu;
        if (fabs(cosdph) > 1.0) cosdph = (cosdph > 1.0) ? 1.0 : -1.0;
        phi = phip - acos(cosdph);
        if (phi > R2P) phi = phi - R2P;
        if (phi < 0.0) phi = phi + R2P;
        sin2i2 = 2.0 * sini2 * cosi2; cos2i2 = 2.0 * (cosi2 * cosi2) - 1.0;
        sin2 = sin2i2 * sin2i1; cos2 = cos2i2 * cos2i1;
        sin2cos1 = sin2i2 * cos2i1; cos2sin1 = cos2i2 * sin2i1;
        a21 = p[1] * cos2i2; a22 = p[0] * cos2 - p[2] * sin2; a23 = -p[0] * cos2sin1 - p[2] * sin2cos1; a24 = p[3] * sin2i2;
        a31 = p[1] * sin2i2; a32 = p[0] * sin2cos1 + p[2] * cos2sin1; a33 = -p[0] * sin2 + p[2] * cos2; a34 = -p[3] * cos2i2;
        a42 = p[3] * sin2i1; a43 = p[3] * cos2i1; a44 = p[2];
    }
    s[0] = (a11 * sip + a12 * sqp + a13 * sup) / a;
    s[1] = (a21 * sip + a22 * sqp + a23 * sup + a24 * svp) / a;
    s[2] = (a31 * sip + a32 * sqp + a33 * sup + a34 * svp) / a;
    s[3] = (a42 * sqp + a43 * sup + a44 * svp) / a;
    dir[0] = cost; dir[1] = sint; dir[2] = phi; dir[3] = cos(phi); dir[4] = sin(phi);
}

/* stokespeel.f90:77-322 (no division by the phase function) */
void orc_stokespeel(double g, double pl, double s[4], double cost, double sint,
                    double ri1, double bmu, double costp, double sintp)
{
    double sip = s[0], sqp = s[1], sup = s[2], svp = s[3];
    double p[4], cosb2 = bmu * bmu, sinbt;
    double a11, a12, a13, a21, a22, a23, a24, a31, a32, a33, a34, a42, a43, a44;
    double sini2 = 0.0, cosi2 = 0.0, bott, sin2i2, cos2i2, sin2, cos2, sin2cos1, cos2sin1;
    orc_dustmat(g, pl, bmu, cosb2, p);
    sinbt = sqrt(1.0 - cosb2);
    if (ri1 > PI) {
        double ri3 = R2P - ri1, cosi3 = cos(ri3), sini3 = sin(ri3);
        double sin2i3 = 2.0 * sini3 * cosi3, cos2i3 = 2.0 * (cosi3 * cosi3) - 1.0;
        a11 = p[0]; a12 = p[1] * cos2i3; a13 = p[1] * sin2i3;
        if (bmu == 1.0) return;
        else if (bmu == -1.0) { sini3 = 1.0; sini2 = 1.0; cosi3 = 0.0; cosi2 = 0.0; }
        else {
            if (fabs(cost) < 1.0) {
                sini2 = sini3 * sintp / sint;
                bott = sint * sinbt;
                cosi2 = costp / bott - cost * bmu / bott;
            } else {
                sini2 = 0.0;
                if (cost >= 1.0) cosi2 = -1.0;
                if (cost <= -1.0) cosi2 = 1.0;
            }
        }
        sin2i2 = 2.0 * sini2 * cosi2; cos2i2 = 2.0 * (cosi2 * cosi2) - 1.0;
        sin2 = sin2i2 * sin2i3; cos2 = cos2i2 * cos2i3;
        sin2cos1 = sin2i2 * cos2i3; cos2sin1 = cos2i2 * sin2i3;
        a21 = p[1] * cos2i2; a22 = p[0] * cos2 - p[2] * sin2; a23 = p[0] * cos2sin1 + p[2] * sin2cos1; a24 = -p[3] * sin2i2;
        a31 = -p[1] * sin2i2; a32 = -p[0] * sin2cos1 - p[2] * cos2sin1; a33 = -p[0] * sin2 + p[2] * cos2; a34 = -p[3] * cos2i2;
        a42 = -p[3] * sin2i3; a43 = p[3] * cos2i3; a44 = p[2];
    } else {
        double cosi1 = cos(ri1), sini1 = sin(ri1);
        double sin2i1 = 2.0 * sini1 * cosi1, cos2i1 = 2.0 * (cosi1 * cosi1) - 1.0;
        a11 = p[0]; a12 = p[1] * cos2i1; a13 = -p[1] * sin2i1;
        if (bmu == 1.0) return;
        else if (bmu == -1.0) { sini1 = 1.0; sini2 = 1.0; cosi1 = 0.0; cosi2 = 0.0; }
        else {
            if (fabs(cost) < 1.0) {
                sini2 = sini1 * sintp / sint;
                bott = sint * sinbt;
                cosi2 = costp / bott - cost * bmu / bott;
            } else {
                sini2 = 0.0;
                if (cost >= 1.0) cosi2 = -1.0;
                if (cost <= -1.0) cosi2 = 1.0;
            }
        }
        sin2i2 = 2.0 * sini2 * cosi2; cos2i2 = 2.0 * (cosi2 * cosi2) - 1.0;
        sin2 = sin2i2 * sin2i1; cos2 = cos2i2 * cos2i1;
        sin2cos1 = sin2i2 * cos2i1; cos2sin1 = cos2i2 * sin2i1;
        a21 = p[1] * cos2i2; a22 = p[0] * cos2 - p[2] * sin2; a23 = -p[0] * cos2sin1 - p[2] * sin2cos1; a24 = p[3] * sin2i2;
        a31 = p[1] * sin2i2; a32 = p[0] * sin2cos1 + p[2] * cos2sin1; a33 = -p[0] * sin2 + p[2] * cos2; a34 = -p[3] * cos2i2;
        a42 = p[3] * sin2i1; a43 = p[3] * cos2i1; a44 = p[2];
    }
    s[0] = (a11 * sip + a12 * sqp + a13 * sup);
    s[1] = (a21 * sip + a22 * sqp + a23 * sup + a24 * svp);
    s[2] = (a31 * sip + a32 * sqp + a33 * sup + a34 * svp);
    s[3] = (a42 * sqp + a43 * sup + a44 * svp);
}

/* ======================================================================
 * Frequencies and dust
 * ====================================================================== */

/* random.f90:74-92 zetainit: partial sums of zeta(4), added smallest term first */
static double inczeta[145];
static int inczeta_ready = 0;
static void zetainit(void)
{
    if (inczeta_ready) return;
    for (int i = 0; i < 145; ++i) inczeta[i] = 0.0;
    for (int j = 145; j >= 1; --j) {
        double temp = 1.0 / (double)(j * j * j * j);
        for (int i = 145; i >= j; --i) inczeta[i - 1] += temp;
    }
    inczeta_ready = 1;
}

/* random.f90:44-72 bbfreq: Planck sample x = -ln(z1 z2 z3 z4)/i, i ~ 1/i^4 */
static double bbfreq(orc_rng *r, double teff)
{
    zetainit();
    double z0 = inczeta[144] * ran(r);
    double z1 = ran(r), z2 = ran(r), z3 = ran(r), z4 = ran(r);
    int i = 1;
    while (z0 > inczeta[i - 1]) i++;
    return -log(z1 * z2 * z3 * z4) * teff / (double)i;
}

double orc_bbfreq_lecuyer(double teff)
{
    orc_rng r = { 1, 0, 0, 0 };
    return bbfreq(&r, teff);
}

/* dust_5.f90:9-109 opacset for dust species (no gas mixing) */
typedef struct { double kappa, albedo, g, pl; } orc_opac;
static void opacset(const orc_model *m, double nu, orc_opac *o)
{
    for (int id = 0; id < m->nopac; ++id) {
        const double *nd = m->nudust + (size_t)id * m->nlam;
        const double *kd = m->kappad + (size_t)id * m->nlam, *ad = m->adust + (size_t)id * m->nlam;
        const double *gd = m->gdust + (size_t)id * m->nlam, *pd = m->pldust + (size_t)id * m->nlam;
        int nl = m->nlam;
        if (nu >= nd[0]) { o[id].g = gd[0]; o[id].pl = pd[0]; o[id].kappa = kd[0]; o[id].albedo = ad[0]; }
        else if (nu <= nd[nl - 1]) { o[id].g = gd[nl - 1]; o[id].pl = pd[nl - 1]; o[id].kappa = kd[nl - 1]; o[id].albedo = ad[nl - 1]; }
        else {
            int i0 = 1, i1 = nl, i;
            while (i1 - i0 > 1) {
                i = (i1 + i0) / 2;
                if (nu <= nd[i - 1]) i0 = i; else i1 = i;
            }
            double f = (nu - nd[i1 - 1]) / (nd[i0 - 1] - nd[i1 - 1]);
            double fm1 = 1.0 - f;
            o[id].g = fm1 * gd[i1 - 1] + f * gd[i0 - 1];
            o[id].pl = fm1 * pd[i1 - 1] + f * pd[i0 - 1];
            o[id].kappa = fm1 * kd[i1 - 1] + f * kd[i0 - 1];
            o[id].albedo = fm1 * ad[i1 - 1] + f * ad[i0 - 1];
        }
    }
}

/* lucydustfreq_2_test.f90:19-160 with the uniform xi supplied */
double orc_lucydustfreq(const orc_model *m, double T0, int32_t iop, double xi)
{
    int nT = m->nT, nnu = m->nnu;
    const double *lT = m->lTint, *nui = m->nuint;
    const double *K = m->kapint + (size_t)iop * nT * nnu;
#define KAP(i, j) K[(size_t)((j) - 1) * nT + ((i) - 1)]
    double lT0 = log(T0);
    int j = (int)((double)(nT - 1) * (lT0 - lT[0]) / (lT[nT - 1] - lT[0])) + 1;
    if (j < 1) j = 1;                 /* reference only warns; clamp keeps reads in bounds */
    if (j > nT - 1) j = nT - 1;
    double f = (lT0 - lT[j - 1]) / (lT[j] - lT[j - 1]);
    if (f < 0.0) f = 0.0;
    if (f > 1.0) f = 1.0;
    int i0 = 1, i1 = nnu, i;
    double P0 = (1.0 - f) * KAP(j, i0) + f * KAP(j + 1, i0), P1, P = 0.0;
    if (xi < P0) {
        if (P0 > 0.0) return T0 * xi * nui[0] / P0;
        return T0 * nui[0];
    }
    P1 = (1.0 - f) * KAP(j, i1) + f * KAP(j + 1, i1);
    if (P1 > 1.0) P1 = 1.0;
    if (xi >= P1) {
        if (P1 < 1.0) {
            if (xi < 1.0) return T0 * nui[nnu - 1] * (1.0 - P1) / (1.0 - xi);
            return T0 * nui[nnu - 1];
        }
        return T0 * nui[nnu - 1];
    }
    while (i1 - i0 > 1) {
        i = (i0 + i1) / 2;
        P = (1.0 - f) * KAP(j, i) + f * KAP(j + 1, i);
        if (P > 1.0) P = 1.0;
        if (P < P0) { P0 = P; i0 = i; }
        if (P > P1) { P1 = P; i1 = 1; }   /* reference sets i1=1 here (lucydustfreq_2_test.f90:112) */
        if (P > xi) { i1 = i; P1 = P; } else { i0 = i; P0 = P; }
    }
    if (xi < P0) f = 0.0;
    else if (xi > P1) f = 1.0;
    else if (P0 <= 0.0) f = 0.5;
    else if (P1 <= 0.0) { /* reference keeps f */ }
    else if (xi <= 0.0) { /* reference keeps f */ }
    else {
        /* P1/P0 rounds to 1 only when P0, P1 are adjacent doubles; the
         * reference would divide 0 by 0 there, here f = 0 */
        double den = log(P1 / P0);
        f = den > 0.0 ? log(xi / P0) / den : 0.0;
    }
    return T0 * nui[i0 - 1] * pow(nui[i1 - 1] / nui[i0 - 1], f);
#undef KAP
}

static double lucydustfreq(const orc_model *m, orc_rng *r, double T0, int32_t iop)
{
    if (T0 >= TMAXK / 11605.0) return bbfreq(r, T0);
    double xi = ran(r);
    return orc_lucydustfreq(m, T0, iop, xi);
}

/* dust_5.f90:144-172 KapP */
static double kapP(const orc_model *m, double T, int32_t iop)
{
    int nT = m->nT;
    double lT = log(T);
    int k = (int)((double)(nT - 1) * (lT - m->lTint[0]) / (m->lTint[nT - 1] - m->lTint[0])) + 1;
    const double *kp = m->kappap + (size_t)iop * nT;
    if (k < 1) return kp[0];
    if (k > nT - 1) return kp[nT - 1];
    double f = (lT - m->lTint[k - 1]) / (m->lTint[k] - m->lTint[k - 1]);
    return (1.0 - f) * kp[k - 1] + f * kp[k];
}

static double pow4(double x) { double x2 = x * x; return x2 * x2; }

/* dusttemp_5.f90:8-152 with zero velocities (adi = adv = 0) */
double orc_dusttemp(const orc_model *m, double na, double nphot, int32_t iop, double mcell)
{
    const double Tfloor = 0.1 / 11605.0, Tcap = TMAXK / 11605.0;
    int nT = m->nT;
    const double *kp = m->kappap + (size_t)iop * nT;
    double Td = Tfloor;
    (void)kapP;
    if (!(mcell > 0.0)) return Tfloor;
    double absv = pow4(m->tstar) * (na * PI * m->rstar * m->rstar) / (nphot * m->kappav[iop]);
    absv = absv * (RSOL * RSOL) / pow4(mcell);
    if (na == 0.0) return Tfloor;
    double minkap = pow4(m->Tint[0]) * kp[0], maxkap = minkap;
    int mini = 1, maxi = 1;
    for (int i = 1; i <= nT; ++i) {
        double kt = pow4(m->Tint[i - 1]) * kp[i - 1];
        if (minkap > kt) { minkap = kt; mini = i; }
        if (maxkap < kt) { maxkap = kt; maxi = i; }
    }
    if (maxkap < absv) { Td = m->Tint[maxi - 1]; if (Td > Tcap) Td = Tcap; if (Td < Tfloor) Td = Tfloor; }
    if (minkap > absv) { Td = m->Tint[mini - 1]; if (Td > Tcap) Td = Tcap; if (Td < Tfloor) Td = Tfloor; }
    if (mini < maxi) {
        for (int i = mini; i <= maxi - 1; ++i) {
            double kl = pow4(m->Tint[i - 1]) * kp[i - 1], ku = pow4(m->Tint[i]) * kp[i];
            if (kl < absv && ku >= absv) {
                double Tl = m->Tint[i - 1], Tu = m->Tint[i];
                Td = Tl + (Tu - Tl) / (ku - kl) * (absv - kl);
                if (Td > Tcap) Td = Tcap;
                if (Td < Tfloor) Td = Tfloor;
                break;
            }
        }
    } else {
        for (int i = maxi; i <= mini - 1; ++i) {
            double kl = pow4(m->Tint[i - 1]) * kp[i - 1], ku = pow4(m->Tint[i]) * kp[i];
            if (kl > absv && ku <= absv) {
                double Tl = m->Tint[i - 1], Tu = m->Tint[i];
                Td = Tl + (Tu - Tl) / (ku - kl) * (absv - kl);
                if (Td > Tcap) Td = Tcap;
                if (Td < Tfloor) Td = Tfloor;
                break;
            }
        }
    }
    return Td;
}

/* tfinal_5_test1.f90:206-281 (one pass; velocities zero so the smoothing
 * and gradient terms of :61-205 do not change T): Lucy temperature over the
 * upper hemisphere at ip=1 averaged with its mirror cell, mirrored to the
 * lower hemisphere and replicated in phi. */
static int orc_diffusing(const orc_model *m, int ir, int it)
{
    if (!m->diff_itmin) return 0;
    int a = m->diff_itmin[ir - 1];
    return a > 0 && it >= a && it <= m->ntg - a;
}

void orc_tfinal2(const orc_model *m, const double *dtauabs, double nphot, const double *tdust, double *tdave)
{
    const double Tfloor = 0.1 / 11605.0;
    int nrg = m->nrg, ntg = m->ntg, npg = m->npg, ip = 1;
    for (int ir = 1; ir <= nrg - 1; ++ir)
        for (int it = 1; it <= (ntg - 1) / 2; ++it) {
            size_t c = CELL(m, ir, it, ip), cm = CELL(m, ir, ntg - it, ip);
            double t = Tfloor;
            int sp = m->species[c];
            if (sp > 0 && m->rho[c] > 0.0 && tdust && orc_diffusing(m, ir, it)) {
                t = (tdust[c] + tdust[cm]) / 2.0;           /* tfinal_5_test1.f90:232-233 */
                if (t * 11605.0 >= TMAXK) t = TMAXK / 11605.0;
            } else if (sp > 0 && m->rho[c] > 0.0) {
                int iop = m->species_opac[sp];
                double rho2 = (m->species[cm] == sp) ? m->rho[cm] : 0.0;
                double mcell1 = pow(m->rho[c] * m->vcell[c], 0.25);
                double mcell2 = pow(rho2 * m->vcell[c], 0.25);
                double t1 = orc_dusttemp(m, dtauabs[c], nphot, iop, mcell1);
                double t2 = orc_dusttemp(m, dtauabs[cm], nphot, iop, mcell2);
                t = (t1 + t2) / 2.0;
                if (t * 11605.0 >= TMAXK) t = TMAXK / 11605.0;
            }
            tdave[c] = t;
        }
    for (int ir = 1; ir <= nrg - 1; ++ir)
        for (int it = (ntg + 1) / 2; it <= ntg - 1; ++it)
            tdave[CELL(m, ir, it, 1)] = tdave[CELL(m, ir, ntg - it, 1)];
    if (npg > 2)
        for (int ipp = 2; ipp <= npg - 1; ++ipp)
            for (int it = 1; it <= ntg - 1; ++it)
                for (int ir = 1; ir <= nrg - 1; ++ir)
                    tdave[CELL(m, ir, it, ipp)] = tdave[CELL(m, ir, it, 1)];
}

void orc_tfinal(const orc_model *m, const double *dtauabs, double nphot, double *tdave)
{
    orc_tfinal2(m, dtauabs, nphot, NULL, tdave);
}

/* chiR, dust_5.f90:112-141 (the index clamped into the table instead of the stop) */
static double orc_chiR(const orc_model *m, const double *kappar, double T, int iop)
{
    int nT = m->nT;
    double lT = log(T);
    int k = (int)((double)(nT - 1) * (lT - m->lTint[0]) / (m->lTint[nT - 1] - m->lTint[0])) + 1;
    if (k < 1) k = 1;
    if (k > nT - 1) k = nT - 1;
    double f = (lT - m->lTint[k - 1]) / (m->lTint[k] - m->lTint[k - 1]);
    return (1.0 - f) * kappar[(size_t)iop * nT + k - 1] + f * kappar[(size_t)iop * nT + k];
}

/* setdiffus_5.f90:22-103 as written: for every column ir from nr0 while
 * rarr(ir) < rmaxd and every it from (ntg-1)/2 down to 2, the Rosseland
 * optical depth toward the pole and the mean free path at depth 1; cells
 * it..(ntg-1)/2 diffuse (diffdir -2) when the depth reaches 1 within
 * hdisk/20; mirrored (+2) and copied to every phi. */
void orc_setdiffus(const orc_model *m, const double *kappar, const double *hdisk, const double *tdave,
                   double rmaxd, int32_t nr0, int32_t iter, uint8_t *diffus, int8_t *diffdir, int32_t *itmin)
{
    const int nrg = m->nrg, ntg = m->ntg, npg = m->npg;
    const size_t nc = (size_t)(nrg - 1) * (ntg - 1) * (npg - 1);
    memset(diffus, 0, nc);
    memset(diffdir, 0, nc);
    const double kapd_cm = m->rstar * RSOL;           /* kapd = kappav * rmincgs (gridset:79) */
    int ir = nr0;
    double rad = m->rarr[ir - 1];
    while (rad < rmaxd && ir <= nrg - 1) {
        double zmr = hdisk[ir - 1];
        for (int it = (ntg - 1) / 2; it >= 2; --it) {
            double tau = 0.0, tau1 = 0.0;
            int itt = it;
            while (tau < 1.0 && itt > 1) {
                double dt = m->thetarr[itt] - m->thetarr[itt - 1];
                tau1 = tau;
                size_t c = CELL(m, ir, itt, 1);
                double tt = tdave[c];
                if (iter < 2) tt = fmax(tt, 50.0 / 11605.0);
                int sp = m->species[c];
                if (sp > 0 && m->rho[c] > 0.0) {
                    int iop = m->species_opac[sp];
                    tau = tau + m->kappav[iop] * kapd_cm * orc_chiR(m, kappar, tt, iop) * m->rho[c] * dt * rad;
                }
                itt = itt - 1;
            }
            double lmfp1;
            if (tau > 1.0) lmfp1 = rad * (m->thetarr[itt + 1] - m->thetarr[itt]) / (tau - tau1) * (1.0 - tau1);
            else lmfp1 = rad * (m->thetarr[itt + 1] - m->thetarr[itt]);
            double lmfp = rad * (m->thetarr[it] - m->thetarr[itt + 1]) + lmfp1;
            if (tau >= 1.0 && lmfp < zmr / 20.0)
                for (int it1 = it; it1 <= (ntg - 1) / 2; ++it1) {
                    diffus[CELL(m, ir, it1, 1)] = 1;
                    diffdir[CELL(m, ir, it1, 1)] = -2;
                }
        }
        ir = ir + 1;
        rad = m->rarr[ir - 1];
    }
    for (int r = 1; r <= nrg - 1; ++r)
        for (int it = (ntg + 1) / 2; it <= ntg - 1; ++it) {
            diffus[CELL(m, r, it, 1)] = diffus[CELL(m, r, ntg - it, 1)];
            diffdir[CELL(m, r, it, 1)] = (int8_t)-diffdir[CELL(m, r, ntg - it, 1)];
        }
    if (npg > 2)
        for (int ip = 2; ip <= npg - 1; ++ip)
            for (int it = 1; it <= ntg - 1; ++it)
                for (int r = 1; r <= nrg - 1; ++r) {
                    diffus[CELL(m, r, it, ip)] = diffus[CELL(m, r, it, 1)];
                    diffdir[CELL(m, r, it, ip)] = diffdir[CELL(m, r, it, 1)];
                }
    for (int r = 1; r <= nrg - 1; ++r) {
        itmin[r - 1] = 0;
        for (int it = 1; it <= (ntg - 1) / 2; ++it)
            if (diffus[CELL(m, r, it, 1)]) { itmin[r - 1] = it; break; }
    }
}

/* emit_common_2.f90:256-341, the state a column's last diffusion photon
 * leaves: the surface T^4 = T0^4 + N delta/dA, the interior integrated down
 * to the midplane with the accretion photons below (nabs). */
void orc_diff_columns(const orc_model *m, const double *kappar, const uint32_t *count, const int32_t *nabs,
                      double delta, double *tdust)
{
    const int nrg = m->nrg, ntg = m->ntg, npg = m->npg;
    const double Tcap = TMAXK / 11605.0, T0 = 0.1 / 11605.0;
    for (int ip = 1; ip <= npg - 1; ++ip)
        for (int ir = 1; ir <= nrg - 1; ++ir)
            for (int h = 0; h < 2; ++h) {
                int a = m->diff_itmin[ir - 1];
                if (a <= 0) continue;
                int upper = (h == 0);
                int its = upper ? a : ntg - a, mid = upper ? (ntg - 1) / 2 : (ntg + 1) / 2, step = upper ? 1 : -1;
                int w = upper ? its : its + 1;
                double sth = sin(m->thetarr[w - 1]);
                double dA = PI * sth * (m->rarr[ir] * m->rarr[ir] - m->rarr[ir - 1] * m->rarr[ir - 1]);
                size_t cs = CELL(m, ir, its, ip);
                double T4 = T0 * T0 * T0 * T0 + (double)count[cs] * delta / dA;
                double Ts = pow(T4, 0.25);
                if (Ts > Tcap) Ts = Tcap;
                if (!(m->species[cs] > 0 && m->rho[cs] > 0.0)) Ts = T0;
                tdust[cs] = Ts;
                double nacc = 0.0;
                for (int i = its; upper ? i <= mid : i >= mid; i += step) nacc += (double)nabs[CELL(m, ir, i, ip)];
                double T4d = T4;
                double rc = 0.5 * (m->rarr[ir - 1] + m->rarr[ir]);
                for (int i = its + step; upper ? i <= mid : i >= mid; i += step) {
                    int ib = i - step;
                    size_t cb = CELL(m, ir, ib, ip);
                    int sp = m->species[cb];
                    if (sp > 0 && m->rho[cb] > 0.0) {
                        int iop = m->species_opac[sp];
                        double dth = upper ? m->thetarr[i - 1] - m->thetarr[i - 2] : m->thetarr[i + 1] - m->thetarr[i];
                        T4d = T4d + 0.75 * delta / dA * nacc * m->kappav[iop] * RSOL * m->rstar * m->rho[cb] *
                                        orc_chiR(m, kappar, tdust[cb], iop) * rc * dth;
                        double Ti = pow(T4d, 0.25);
                        tdust[CELL(m, ir, i, ip)] = Ti > Tcap ? Tcap : Ti;
                        nacc = nacc - (double)nabs[cb];
                    }
                }
            }
}

/* tfinal_5_test1.f90:334-368: the convergence statistic.  Over the upper
 * hemisphere at ip = 1, cells whose old and new temperatures both exceed
 * the 0.1 K floor and differ give x = (tnew - told)/tnew; out[0] = sum |x|,
 * out[1] = sum x^2, out[2] = count; returns sig = sqrt(<x^2> - <|x|>^2). */
double orc_lucy_sigma(const orc_model *m, const double *told, const double *tnew, double out[3])
{
    const double floor_t = 0.1 / 11605.0;
    double a = 0.0, b = 0.0, n = 0.0;
    for (int it = 1; it <= (m->ntg - 1) / 2; ++it)
        for (int ir = 1; ir <= m->nrg - 1; ++ir) {
            size_t c = CELL(m, ir, it, 1);
            double to = told[c], tn = tnew[c];
            if (tn != to && tn > floor_t && to > floor_t) {
                double x = (tn - to) / tn;
                a = fabs(x) + a;
                b = x * x + b;
                n = n + 1.0;
            }
        }
    out[0] = a; out[1] = b; out[2] = n;
    return sqrt(b / n - (a / n) * (a / n));
}

/* ======================================================================
 * Packet state and the cell walk
 * ====================================================================== */
typedef struct {
    double xp, yp, zp, rsq, rtot;
    double ux, uy, uz;
    double cost, sint, phi, cosp, sinp;
    double sip, sqp, sup, svp;
    double nub, wave;
    double cosb, sinb, lp;
    int32_t ir, it, ip, on_wall;
    int32_t iflag, i_orig, exitflag;
    double kapd[8];
    orc_opac op[8];
    double hit, htot;
} orc_pkt;

typedef struct {
    const orc_model *m;
    const double *tdust;
    orc_tallies *t;
    orc_rng rng;
} orc_run;

static void set_kapd(orc_pkt *p, const orc_model *m, int order)
{   /* kapd(id)=kappa(id)*R*[cm]*kappav(id) in the three orders of the call sites */
    for (int id = 0; id < m->nopac; ++id) {
        if (order == 0) p->kapd[id] = p->op[id].kappa * (m->rstar * RSOL) * m->kappav[id];      /* newdisktherm: rmincgs */
        else if (order == 1) p->kapd[id] = p->op[id].kappa * m->rstar * RSOL * m->kappav[id];   /* propagate:157-159 */
        else p->kapd[id] = p->op[id].kappa * RSOL * m->rstar * m->kappav[id];                   /* tauint:199-201 */
    }
}

/* initp_common.f90:1-31 (limb = 0) */
static void initp_sample(orc_run *R, orc_pkt *p)
{
    double xran = ran(&R->rng);
    p->cost = sqrt(xran);
    p->sint = sqrt(1.0 - p->cost * p->cost);
    xran = ran(&R->rng);
    p->phi = R2P * xran;
    p->cosp = cos(p->phi);
    p->sinp = sin(p->phi);
}

/* initp_common.f90:31-92 */
static void initp_transform(orc_pkt *p)
{
    double costi, sinti, cospn, phinew;
    if (fabs(p->cosb) < 0.9999999) {
        costi = p->cost * p->cosb - p->sint * p->sinb * p->cosp;
        if (fabs(costi) > 1.0) { if (costi > 1.0) costi = 1.0; if (p->cost < -1.0) costi = -1.0; }
        sinti = sqrt(1.0 - costi * costi);
        if (sinti > 0.0000001) {
            cospn = (p->cost - costi * p->cosb) / (sinti * p->sinb);
            if (fabs(cospn) > 1.0) cospn = (cospn < 0.0) ? -1.0 : 1.0;
            if (p->phi < PI) { phinew = acos(cospn); if (phinew > PI) phinew = PI; }
            else { phinew = R2P - acos(cospn); if (phinew < PI) phinew = PI; }
        } else phinew = 0.0;
        p->phi = phinew; p->cost = costi; p->sint = sinti;
    } else {
        if (p->cosb < 0.0) p->cost = -p->cost;
    }
    p->phi = p->phi + p->lp;
    if (p->phi > R2P) p->phi = p->phi - R2P;
    if (p->phi < 0.0) p->phi = p->phi + R2P;
    p->sinp = sin(p->phi);
    p->cosp = cos(p->phi);
}

/* testgrid.f90:40-124.  Returns 1 when a reset happened. */
static int testgrid(orc_run *R, orc_pkt *p)
{
    const orc_model *m = R->m;
    const double tol = 1.0e-4;
    int nrg = m->nrg, ntg = m->ntg, npg = m->npg, reset = 0, changed = 0;
    double cosb = check(p->zp / p->rtot);
    double phi = 0.0;
    if (npg > 2) { phi = atan2(p->yp, p->xp); if (phi < 0.0) phi = phi + R2P; }
    const double *r2 = m->r2arr, *ca = m->costarr, *pa = m->phiarr;
    if ((p->rsq - r2[0]) / (r2[1] - r2[0]) < -tol || (p->rsq - r2[nrg - 2]) / (r2[nrg - 1] - r2[nrg - 2]) > 1.0 + tol) {
        p->rtot = pow(ran(&R->rng) * (pow(m->rarr[nrg - 1], 3.0) - pow(m->rarr[0], 3.0)) + pow(m->rarr[0], 3.0), (double)(1.0f / 3.0f));
        reset = 1;
    }
    if ((cosb - ca[0]) / (ca[1] - ca[0]) < -tol || (cosb - ca[ntg - 2]) / (ca[ntg - 1] - ca[ntg - 2]) > 1.0 + tol) {
        cosb = (ran(&R->rng) * (ca[ntg - 1] - ca[0]) + ca[0]);
        reset = 1;
    }
    if (npg > 2 && ((phi - pa[0]) / (pa[1] - pa[0]) < -tol || (phi - pa[npg - 2]) / (pa[npg - 1] - pa[npg - 2]) > 1.0 + tol)) {
        phi = ran(&R->rng) * (pa[npg - 1] - pa[0]) + pa[0];
        reset = 1;
    }
    if (reset) {
        double s = sqrt(1.0 - pow(cosb, 2.0));
        p->xp = p->rtot * s * cos(phi);
        p->yp = p->rtot * s * sin(phi);
        p->zp = p->rtot * cosb;
        p->rsq = p->rtot * p->rtot;
        p->on_wall = 0;
        changed = 1;
    }
    if (p->ir < 1 || p->ir > nrg) { p->ir = orc_locate(r2, nrg, p->rsq); changed = 1; }
    if (p->it < 1 || p->it > ntg) { p->it = orc_locate(ca, ntg, cosb); changed = 1; }
    if (p->ip < 1 || p->ip > npg) { p->ip = orc_locate(pa, npg, phi); changed = 1; }
    double rfrac = (p->rsq - r2[p->ir - 1]) / (r2[p->ir] - r2[p->ir - 1]);
    double tfrac = (cosb - ca[p->it - 1]) / (ca[p->it] - ca[p->it - 1]);
    int phierr = 0;
    if (npg > 2) {
        double pfrac = (phi - pa[p->ip - 1]) / (pa[p->ip] - pa[p->ip - 1]);
        if (pfrac < -tol || pfrac > 1.0 + tol) {
            double d1 = phi - pa[p->ip - 1], d2 = phi - pa[p->ip];
            if (d1 > PI) d1 -= R2P;
            if (d1 < -PI) d1 += R2P;
            if (d2 > PI) d2 -= R2P;
            if (d2 < -PI) d2 += R2P;
            phierr = fabs(d1) > tol && fabs(d2) > tol;
        }
    }
    if (rfrac < -tol || rfrac > 1.0 + tol) { p->ir = orc_locate(r2, nrg, p->rsq); changed = 1; }
    if (tfrac < -tol || tfrac > 1.0 + tol) { p->it = orc_locate(ca, ntg, cosb); changed = 1; }
    if (phierr) { p->ip = orc_locate(pa, npg, phi); changed = 1; }
    return changed;
}

/* one face-selection step shared by tauint and peeloff */
static int next_face(const orc_model *m, const orc_pkt *p, double ux, double uy, double uz, double *dt, int32_t *iface)
{
    orc_walls w;
    orc_reset_t(&w);
    orc_radface(&w, ux, uy, uz, p->xp, p->yp, p->zp, m->r2arr[p->ir - 1], m->r2arr[p->ir]);
    if (m->ntg > 2) orc_thetaface(&w, ux, uy, uz, p->xp, p->yp, p->zp, m->tan2arr[p->it - 1], m->tan2arr[p->it]);
    if (m->npg > 2 && m->ntg > 2)
        orc_phiface(&w, m->aarr[p->ip - 1], m->barr[p->ip - 1], m->aarr[p->ip], m->barr[p->ip], ux, uy, uz, p->xp, p->yp, p->zp);
    if (orc_find_next_wall(&w, p->on_wall, dt, iface) != 0) return -1;
    if (*dt < 0.0) return -1;
    return 0;
}

static void cross_face(const orc_model *m, orc_pkt *p, int32_t iface)
{   /* tauint_3d_2_fuv.f90:148-167 */
    switch (iface) {
    case 1: p->ir--; break;
    case 2: p->ir++; break;
    case 3: p->it--; break;
    case 4: p->it++; break;
    case 5: p->ip--; break;
    default: p->ip++; break;
    }
    if (p->ip == 0) p->ip = m->npg - 1;
    if (p->ip == m->npg) p->ip = 1;
}

static double cell_opac(const orc_model *m, const orc_pkt *p, size_t c, int *iop)
{
    int sp = m->species[c];
    double rho = m->rho[c];
    *iop = -1;
    if (sp > 0 && rho > 0.0) { *iop = m->species_opac[sp]; return 0.0 + p->kapd[*iop] * rho; }
    return 0.0;
}

static void initpabs(orc_run *R, orc_pkt *p);

/* tauint_3d_2_fuv.f90:5-210.  Returns -1 on a walk error. */
static int tauint(orc_run *R, orc_pkt *p, double tau)
{
    const orc_model *m = R->m;
    double tsum = 0.0, dt;
    int32_t iface;
    for (;;) {
        if (testgrid(R, p)) R->t->counters[ORC_C_RESETS] += 1.0;
        if (next_face(m, p, p->ux, p->uy, p->uz, &dt, &iface) != 0) return -1;
        R->t->counters[ORC_C_SEGMENTS] += 1.0;
        size_t c = CELL(m, p->ir, p->it, p->ip);
        int iop;
        double opac = cell_opac(m, p, c, &iop);
        double dtau = opac * dt;
        if ((tsum + dtau) > tau) {
            double ds = (tau - tsum) / opac;
            if (iop >= 0) R->t->dtauabs[c] += ds * p->kapd[iop] * m->rho[c] * (1.0 - p->op[iop].albedo);
            p->xp = p->xp + ds * p->ux;
            p->yp = p->yp + ds * p->uy;
            p->zp = p->zp + ds * p->uz;
            p->rsq = p->xp * p->xp + p->yp * p->yp + p->zp * p->zp;
            p->rtot = sqrt(p->rsq);
            p->on_wall = 0;
            return 0;
        }
        if (iop >= 0) R->t->dtauabs[c] += dt * p->kapd[iop] * m->rho[c] * (1.0 - p->op[iop].albedo);
        cross_face(m, p, iface);
        p->xp = p->xp + dt * p->ux;
        p->yp = p->yp + dt * p->uy;
        p->zp = p->zp + dt * p->uz;
        p->on_wall = 1;
        p->rsq = p->xp * p->xp + p->yp * p->yp + p->zp * p->zp;
        p->rtot = sqrt(p->rsq);
        tsum = tsum + dtau;
        if (p->ir == m->nrg) { p->exitflag = 1; return 0; }
        if (p->ir == 0) {
            R->t->counters[ORC_C_AFLUX] += 1.0;
            p->rtot = sqrt(p->rsq);
            initpabs(R, p);
            opacset(m, p->nub, p->op);
            p->wave = 1.2398 / p->nub;
            p->i_orig = 1;
            p->iflag = 0;
            set_kapd(p, m, 2);
            p->ir = 1;
            tsum = 0.0;
        }
    }
}

/* initpabs.f90:1-79 (frequency from the blackbody, see header) */
static void initpabs(orc_run *R, orc_pkt *p)
{
    p->sip = 1.0; p->sqp = 0.0; p->sup = 0.0; p->svp = 0.0;
    p->cosb = p->zp / p->rtot;
    p->sinb = sqrt(1.0 - p->cosb * p->cosb);
    p->lp = atan2(p->yp, p->xp);
    initp_sample(R, p);
    initp_transform(p);
    p->ux = p->sint * p->cosp;
    p->uy = p->sint * p->sinp;
    p->uz = p->cost;
    p->nub = bbfreq(&R->rng, R->m->tstar);
}

static int i_orig_of(int sp)
{   /* propagate_5_fuv.f90:123-132 */
    if (sp == 1 || sp == 2 || sp == 5 || sp == 6 || sp == 9) return 2;
    if (sp == 3 || sp == 7) return 3;
    return 4;
}

static int freq_bin(const orc_model *m, double nub)
{   /* propagate_5_fuv.f90:273, peeloff_3d_2_test.f90:289 */
    int inub = (int)((double)m->nfreq * log(nub / m->numin) / m->lnurat) + 1;
    if (inub < 1) inub = 1;
    if (inub > m->nfreq) inub = m->nfreq;
    return inub;
}

/* peeloff_3d_2_test.f90:7-347.  `q` is the caller's packet (copied). */
static void peeloff(orc_run *R, const orc_pkt *q, double sipnew, int peelid, int idust_sp)
{
    const orc_model *m = R->m;
    orc_tallies *t = R->t;
    orc_pkt p = *q;
    double s[4] = { sipnew, q->sqp, q->sup, q->svp };
    double tsum = 0.0, dt;
    int32_t iface;
    int k = peelid - 1;
    double coste = m->coste[k], sinte = m->sinte[k], cospe = m->cospe[k], sinpe = m->sinpe[k];
    double ux1 = sinte * cospe, uy1 = sinte * sinpe, uz1 = coste;
    if (p.on_wall) {
        int inner = fabs(p.rsq - m->r2arr[p.ir - 1]) < fabs(p.rsq - m->r2arr[p.ir]);
        int outwards = ux1 * p.xp + uy1 * p.yp + uz1 * p.zp > 0.0;
        if (inner && !outwards) { if (p.ir == 1) return; p.ir = p.ir - 1; }
        if (!inner && outwards) p.ir = p.ir + 1;
    }
    for (;;) {
        if (testgrid(R, &p)) t->counters[ORC_C_RESETS] += 1.0;
        if (next_face(m, &p, ux1, uy1, uz1, &dt, &iface) != 0) { t->counters[ORC_C_ERRORS] += 1.0; return; }
        t->counters[ORC_C_PEELSEG] += 1.0;
        size_t c = CELL(m, p.ir, p.it, p.ip);
        int iop;
        double opac = cell_opac(m, &p, c, &iop);
        double dtau = opac * dt;
        tsum = tsum + dtau;
        if (tsum > m->taumax) return;
        cross_face(m, &p, iface);
        p.xp = p.xp + dt * ux1;
        p.yp = p.yp + dt * uy1;
        p.zp = p.zp + dt * uz1;
        p.on_wall = 1;
        p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
        p.rtot = sqrt(p.rsq);
        if (p.ir == m->nrg) break;
        if (p.ir == 0) return;
    }
    t->counters[ORC_C_PEELEVT] += 1.0;
    double factor = exp(-tsum);
    s[0] = s[0] * factor; s[1] = s[1] * factor; s[2] = s[2] * factor; s[3] = s[3] * factor;
    if (q->iflag == 1) {
        double bmu, sinbm, phinew, ri1;
        orc_findangle3(q->cost, q->sint, q->phi, q->cosp, q->sinp, coste, sinte, m->phie[k], cospe, sinpe, &bmu, &sinbm, &phinew, &ri1);
        int iop = m->species_opac[idust_sp];
        orc_stokespeel(p.op[iop].g, p.op[iop].pl, s, coste, sinte, ri1, bmu, q->cost, q->sint);
    }
    double yimage = p.zp * sinte - p.yp * coste * sinpe - p.xp * coste * cospe;
    double ximage = p.yp * cospe - p.xp * sinpe;
    double fractxh = 2.0 * m->rmaxi / (double)m->nx;
    int ix = (int)((ximage + m->rmaxi) / fractxh) + 1;
    int iy = (int)((yimage + m->rmaxi) / fractxh) + 1;
    double freq = 2.9979e14 / q->wave;
    if (m->nband > 0 && ix <= m->nx && iy <= m->nx && ix > 0 && iy > 0) {
        size_t npix = (size_t)m->npeel * m->nx * m->nx;
        size_t plane = (size_t)m->nband * npix;
        for (int b = 0; b < m->nband; ++b) {
            const double *fw = m->band_freq + (size_t)b * m->nwmax, *fp = m->band_phi + (size_t)b * m->nwmax;
            int nf = m->band_nw[b] - 1;
            if (freq > fw[0] && freq < fw[nf - 1]) {
                int iw = 1;
                while (freq > fw[iw - 1]) iw++;
                iw = iw - 1;
                double wt = fp[iw - 1];
                size_t o = (size_t)b * npix + (size_t)k + (size_t)m->npeel * ((size_t)(ix - 1) + (size_t)m->nx * (size_t)(iy - 1));
                t->image[0 * plane + o] += (float)(s[0] * wt);
                t->image[1 * plane + o] += (float)(s[1] * wt);
                t->image[2 * plane + o] += (float)(s[2] * wt);
                t->image[3 * plane + o] += (float)(s[3] * wt);
                if (q->i_orig == 1) t->image[5 * plane + o] += (float)(s[0] * wt);
                else if (q->i_orig == 2) t->image[4 * plane + o] += (float)(s[0] * wt);
                else if (q->i_orig == 3) t->image[6 * plane + o] += (float)(s[0] * wt);
                else if (q->i_orig == 4) t->image[7 * plane + o] += (float)(s[0] * wt);
            }
        }
    }
    double rho2 = ximage * ximage + yimage * yimage;
    int inub = freq_bin(m, q->nub);
    int i_inter = (q->iflag == 1) ? 2 : ((q->i_orig == 1) ? 1 : 3);
    int indx[3] = { 1, 1 + q->i_orig, 5 + i_inter };
    size_t nb = (size_t)m->nfreq * m->npeel * m->nap * ORC_NO;
    for (int ia = 1; ia <= m->nap; ++ia) {
        if (rho2 < m->aperture2[ia - 1]) {
            for (int j = 0; j < 3; ++j) {
                size_t o = (size_t)(inub - 1) + (size_t)m->nfreq * ((size_t)k + (size_t)m->npeel * ((size_t)(ia - 1) + (size_t)m->nap * (size_t)(indx[j] - 1)));
                t->psed[0 * nb + o] += (float)s[0];
                t->psed[1 * nb + o] += (float)s[1];
                t->psed[2 * nb + o] += (float)s[2];
                t->psed[3 * nb + o] += (float)s[3];
                t->psed[4 * nb + o] += (float)(s[0] * s[0]);
                t->psed[5 * nb + o] += (float)(s[1] * s[1]);
                t->psed[6 * nb + o] += (float)(s[2] * s[2]);
                t->psed[7 * nb + o] += (float)(s[3] * s[3]);
                t->psed[8 * nb + o] += 1.0f;
            }
        }
    }
}

/* stokes.f90:46-137: rejection sampling of (bmu, ri1), then the update */
static void stokes(orc_run *R, orc_pkt *p, int iop)
{
    double g = p->op[iop].g, pl = p->op[iop].pl;
    double peak = cs_peak(g), bmu, ri1, rprob, pp[4];
    for (;;) {
        p->htot += 1.0;
        bmu = 1.0 - 2.0 * ran(&R->rng);
        double cosb2 = bmu * bmu;
        orc_dustmat(g, pl, bmu, cosb2, pp);
        ri1 = R2P * ran(&R->rng);
        if (ri1 > PI) {
            double ri3 = R2P - ri1, cosi3 = cos(ri3), sini3 = sin(ri3);
            double sin2i3 = 2.0 * sini3 * cosi3, cos2i3 = 2.0 * (cosi3 * cosi3) - 1.0;
            rprob = (pp[0] * p->sip + pp[1] * cos2i3 * p->sqp + pp[1] * sin2i3 * p->sup) / p->sip;
        } else {
            double cosi1 = cos(ri1), sini1 = sin(ri1);
            double sin2i1 = 2.0 * sini1 * cosi1, cos2i1 = 2.0 * (cosi1 * cosi1) - 1.0;
            rprob = (pp[0] * p->sip + pp[1] * cos2i1 * p->sqp + (-pp[1] * sin2i1) * p->sup) / p->sip;
        }
        if (rprob > peak) peak = rprob;
        double xran = ran(&R->rng);
        if (peak * xran > rprob) continue;
        p->hit += 1.0;
        break;
    }
    double dir[5] = { p->cost, p->sint, p->phi, p->cosp, p->sinp };
    double s[4] = { p->sip, p->sqp, p->sup, p->svp };
    orc_stokes_apply(g, pl, bmu, ri1, dir, s);
    p->cost = dir[0]; p->sint = dir[1]; p->phi = dir[2]; p->cosp = dir[3]; p->sinp = dir[4];
    p->sip = s[0]; p->sqp = s[1]; p->sup = s[2]; p->svp = s[3];
}

static int select_dust(orc_run *R, const orc_pkt *p)
{   /* select_dust_2.f90:1-34: one thermal species per cell; the draw is still consumed */
    (void)ran(&R->rng);
    return R->m->species[CELL(R->m, p->ir, p->it, p->ip)];
}

/* reemit.f90:17-34 (lucydustfreq_2_test.f90 for the frequency): isotropic,
 * unpolarized, at the dust temperature tave */
static void reemit(orc_run *R, orc_pkt *p, double tave, int iop)
{
    p->sip = 1.0; p->sqp = 0.0; p->sup = 0.0; p->svp = 0.0;
    p->cost = 1.0 - 2.0 * ran(&R->rng);
    p->sint = sqrt(1.0 - p->cost * p->cost);
    p->phi = R2P * ran(&R->rng);
    p->cosp = cos(p->phi);
    p->sinp = sin(p->phi);
    p->nub = lucydustfreq(R->m, &R->rng, tave, iop);
    p->uy = p->sint * p->sinp; p->ux = p->sint * p->cosp; p->uz = p->cost;
}

/* emit_common_2.f90:226-378 (theta diffusion): a photon absorbed in a
 * diffusion cell moves at fixed radius and azimuth to just outside the
 * region's surface wall, leaves with a cosine law about the outward normal
 * (isotrp.f90) at the surface temperature, which grows by 4 pi T*^4/(ns dA)
 * in T^4 with every photon of that surface cell in the order they arrive
 * (one counter for all threads), plus 0.75 mu T4eff of the accretion
 * photons at the midplane.  The Stokes vector is left as it was. */
static void diffuse(orc_run *R, orc_pkt *p)
{
    const orc_model *m = R->m;
    const int ntg = m->ntg;
    const int a = m->diff_itmin[p->ir - 1];
    const int upper = p->it <= (ntg - 1) / 2;
    const int its = upper ? a : ntg - a;
    const int w = upper ? its : its + 1;
    const double cw = m->costarr[w - 1], sw = sin(m->thetarr[w - 1]);
    double cosb = p->zp / p->rtot;
    double sinb = sqrt(1.0 - cosb * cosb);
    double cosbnew = 1.00001 * cw;
    double sinbnew = sqrt(1.0 - cosbnew * cosbnew);
    p->xp = p->xp * sinbnew / sinb;
    p->yp = p->yp * sinbnew / sinb;
    p->zp = p->rtot * cosbnew;
    double sux, suy, suz;
    if (cw > 0.0) { sux = -cw * p->xp / sw / p->rtot; suy = -cw * p->yp / sw / p->rtot; suz = sw; }
    else { sux = cw * p->xp / sw / p->rtot; suy = cw * p->yp / sw / p->rtot; suz = -sw; }
    double x = 0.0, y = 0.0, z = 0.0, r2 = 2.0, mur = 0.0;
    while (r2 * r2 > fabs(mur)) {
        x = ran(&R->rng); x = x + x - 1.0;
        y = ran(&R->rng); y = y + y - 1.0;
        z = ran(&R->rng); z = z + z - 1.0;
        mur = x * sux + y * suy + z * suz;
        r2 = x * x + y * y + z * z;
    }
    double r = sqrt(r2);
    if (mur < 0.0) r = -r;
    double ux = x / r, uy = y / r, uz = z / r;
    p->cost = uz;
    p->sint = sqrt(1.0 - uz * uz);
    p->cosp = ux / p->sint;
    p->sinp = uy / p->sint;
    p->phi = atan2(p->sinp, p->cosp);
    double dA = PI * sw * (m->rarr[p->ir] * m->rarr[p->ir] - m->rarr[p->ir - 1] * m->rarr[p->ir - 1]);
    size_t cs = CELL(m, p->ir, its, p->ip);
    uint32_t k;
    #pragma omp atomic capture
    k = m->diff_count[cs]++;
    const double Tcap = TMAXK / 11605.0, T0 = 0.1 / 11605.0;
    double Ts = pow(T0 * T0 * T0 * T0 + (double)(k + 1u) * m->diff_delta / dA, 0.25);
    if (Ts > Tcap) Ts = Tcap;
    int mid = upper ? (ntg - 1) / 2 : (ntg + 1) / 2;
    double nacc = (double)R->t->nabs[CELL(m, p->ir, mid, p->ip)];
    double T4eff = m->diff_delta * nacc / dA;
    double mu = ux * sux + uy * suy + uz * suz;
    double Te = pow(Ts * Ts * Ts * Ts + 0.75 * mu * T4eff, 0.25);
    if (Te > Tcap) Te = Tcap;
    p->nub = bbfreq(&R->rng, Te);
    p->it = upper ? its - 1 : its + 1;
    p->ux = p->sint * p->cosp; p->uy = p->sint * p->sinp; p->uz = p->cost;
}

/* propagate_5_fuv.f90:1-326.  Returns exit code: 1 escaped, 2 killed, -1 error. */
static int propagate(orc_run *R, orc_pkt *p, int *n_inter)
{
    const orc_model *m = R->m;
    orc_tallies *t = R->t;
    double xpold = p->xp, ypold = p->yp, zpold = p->zp;
    int iscat = 0, iint = 0, idust2, nint = 0;
    p->iflag = 0; p->exitflag = 0;
    double tau = -log(ran(&R->rng));
    if (tauint(R, p, tau) != 0) return -1;
    if (!p->exitflag) {
        idust2 = select_dust(R, p);
        t->counters[ORC_C_TOT] += 1.0;
        for (;;) {
            size_t c = CELL(m, p->ir, p->it, p->ip);
            int iop = m->species_opac[idust2];
            nint++;
            if (ran(&R->rng) <= p->op[iop].albedo) {
                p->iflag = 1; iscat++; iint++;
                if (m->ipeel == 1)
                    for (int pid = 1; pid <= m->npeel; ++pid) peeloff(R, p, p->sip, pid, idust2);
                stokes(R, p, iop);
            } else {
                t->dbig[c] += 1;
                p->i_orig = i_orig_of(idust2);
                /* propagate_5_fuv.f90:139, emit_common_2.f90:29-100 */
                if (orc_diffusing(m, p->ir, p->it)) diffuse(R, p);
                else {
                    t->nabs[c] += 1;
                    /* emit_common_final.f90:73-81 -> reemit.f90:17-34 */
                    reemit(R, p, R->tdust[c], iop);
                }
                p->iflag = 0;
                opacset(m, p->nub, p->op);
                p->wave = 1.2398 / p->nub;
                set_kapd(p, m, 1);
                if (m->ipeel == 1)
                    for (int pid = 1; pid <= m->npeel; ++pid) peeloff(R, p, p->sip, pid, idust2);
                iint++;
            }
            tau = -log(ran(&R->rng));
            xpold = p->xp; ypold = p->yp; zpold = p->zp;
            p->ux = p->sint * p->cosp; p->uy = p->sint * p->sinp; p->uz = p->cost;
            int pir = p->ir, pit = p->it, pip = p->ip;
            if (tauint(R, p, tau) != 0) return -1;
            if (p->exitflag) break;
            if (p->ir != pir || p->it != pit || p->ip != pip) iint = 0;
            if (iint > m->iintmax) {
                t->killed[CELL(m, p->ir, p->it, p->ip)] += 1;
                t->counters[ORC_C_ABFLUX] += 1.0;
                *n_inter = nint;
                return 2;
            }
            idust2 = select_dust(R, p);
        }
    }
    *n_inter = nint;
    /* SED binning, propagate_5_fuv.f90:231-320 */
    int itb = (int)((double)m->nmu * (1.0 - p->cost) / 2.0) + 1;
    if (p->phi < 0.0) p->phi = p->phi + R2P;
    int ipb = (int)((double)m->nph * p->phi / R2P) + 1;
    if (itb < 1) itb = 1;            /* the reference only warns outside 1..nmu */
    if (itb > m->nmu) itb = m->nmu;
    if (ipb < 1) ipb = 1;
    if (ipb > m->nph) ipb = m->nph;
    double x = ypold * p->cosp - xpold * p->sinp;
    double y = zpold * m->sini_mu[itb - 1] - ypold * m->u_mu[itb - 1] * p->sinp - xpold * m->u_mu[itb - 1] * p->cosp;
    double rho2 = (x * x + y * y);
    if (rho2 < m->aperture2[m->nap - 1] * 1.0001) {
        t->counters[ORC_C_FLUX] += 1.0;
        if (p->iflag == 1) { t->counters[ORC_C_SFLUX] += 1.0; t->counters[ORC_C_NSCAT] += iscat; }
    }
    int inub = freq_bin(m, p->nub);
    int i_inter = (p->iflag == 1) ? 2 : ((p->i_orig == 1) ? 1 : 3);
    int indx[3] = { 1, 1 + p->i_orig, 5 + i_inter };
    size_t nb = (size_t)m->nfreq * m->nmu * m->nph * m->nap * ORC_NO;
    for (int ia = 1; ia <= m->nap; ++ia) {
        if (rho2 < m->aperture2[ia - 1]) {
            for (int j = 0; j < 3; ++j) {
                size_t o = (size_t)(inub - 1) + (size_t)m->nfreq * ((size_t)(itb - 1) + (size_t)m->nmu * ((size_t)(ipb - 1) + (size_t)m->nph * ((size_t)(ia - 1) + (size_t)m->nap * (size_t)(indx[j] - 1))));
                t->sed[0 * nb + o] += (float)p->sip;
                t->sed[1 * nb + o] += (float)p->sqp;
                t->sed[2 * nb + o] += (float)p->sup;
                t->sed[3 * nb + o] += (float)p->svp;
                t->sed[4 * nb + o] += (float)(p->sip * p->sip);
                t->sed[5 * nb + o] += (float)(p->sqp * p->sqp);
                t->sed[6 * nb + o] += (float)(p->sup * p->sup);
                t->sed[7 * nb + o] += (float)(p->svp * p->svp);
                t->sed[8 * nb + o] += 1.0f;
                t->sed[9 * nb + o] += (float)fabs(p->cost);
            }
        }
    }
    if (p->iflag == 1) t->counters[ORC_C_SCOUNT] += 1.0;
    return 1;
}

/* rep_isrf.f90:13-29: ISRF frequency (eV) from the tabulated CDF */
static double rep_isrf(orc_run *R)
{
    const orc_model *m = R->m;
    const double xran = ran(&R->rng);
    int32_t ifreq = orc_locate(m->isrf_cdf, m->nisrf, xran);
    if (ifreq == m->nisrf) ifreq = ifreq - 1;
    if (ifreq == 0) ifreq = 1;
    const double *lf = m->isrf_logfreq - 1, *lc = m->isrf_logcdf - 1;     /* 1-based */
    const double logfreq = lf[ifreq] + (log10(xran) - lc[ifreq]) / (lc[ifreq + 1] - lc[ifreq]) * (lf[ifreq + 1] - lf[ifreq]);
    return pow(10.0, logfreq);
}

/* newdisktherm_5_fuv.f90:416-455 with initpout.f90:18-57: a photon of the
 * external radiation field entering at 0.999999 rmax, directed inward in
 * the local frame, not peeled (the reference subtracts the sky) */
static void emit_outside(orc_run *R, orc_pkt *p)
{
    const orc_model *m = R->m;
    p->sip = 1.0; p->sqp = 0.0; p->sup = 0.0; p->svp = 0.0;
    double xran = ran(&R->rng);
    p->cosb = 1.0 - 2.0 * xran;
    p->sinb = sqrt(1.0 - p->cosb * p->cosb);
    xran = ran(&R->rng);
    p->cost = -xran;
    p->sint = sqrt(1.0 - p->cost * p->cost);
    xran = ran(&R->rng);
    p->phi = R2P * xran;
    p->cosp = cos(p->phi);
    p->sinp = sin(p->phi);
    xran = ran(&R->rng);
    p->lp = R2P * xran;
    initp_transform(p);
    p->nub = rep_isrf(R);
    p->on_wall = 1;
    p->i_orig = 5;
    opacset(m, p->nub, p->op);
    p->wave = 1.2398 / p->nub;
    set_kapd(p, m, 0);
    const double coslp = cos(p->lp), sinlp = sin(p->lp);
    const double rmax = m->rarr[m->nrg - 1];
    const double shrink = (double)0.999999f;        /* a default-real literal in the reference */
    p->zp = p->cosb * rmax * shrink;
    const double rp = p->sinb * rmax * shrink;
    p->xp = rp * coslp;
    p->yp = rp * sinlp;
    p->rsq = p->zp * p->zp + rp * rp;
    p->rtot = sqrt(p->rsq);
    p->ux = p->sint * p->cosp; p->uy = p->sint * p->sinp; p->uz = p->cost;
    p->iflag = 0;
    p->ir = m->nrg - 1;
    p->it = orc_locate(m->thetarr, m->ntg, acos(p->cosb));
    p->ip = orc_locate(m->phiarr, m->npg, p->lp);
}

#define ORC_ACC_TRIES 65536

/* initpacc_4.f90:25-100 with emit_common_2.f90:37-85 (no diffusion) and
 * reemit.f90, then newdisktherm_5_fuv.f90:249-283: a disk-accretion photon.
 * Radial cell j with Ldint(j) <= xi < Ldint(j+1) (locate), the point at
 * ravearr(j) = (rarr(j) + rarr(j+1))/2, azimuth from rantheta
 * (random.f90:23-42), height hdisk(j) gasdev (random.f90:119-138, its saved
 * second deviate kept per packet: the reference keeps one across packets),
 * drawn again at the same radius until the cell holds dust; a point beyond
 * the outer wall is drawn again too (the reference would read past the
 * grid), and after ORC_ACC_TRIES draws the packet is dropped (0).  The dust
 * absorbs it and re-emits it at the cell temperature; peeled as an emission
 * of origin 2. */
static int emit_acc(orc_run *R, orc_pkt *p, int *idust2)
{
    const orc_model *m = R->m;
    orc_tallies *t = R->t;
    const double xi = ran(&R->rng);
    int j = orc_locate(m->acc_ldint, m->nrg - 1, xi);
    if (j < 1) j = 1;
    if (j > m->nrg - 1) j = m->nrg - 1;
    const double rp = 0.5 * (m->rarr[j - 1] + m->rarr[j]);
    const double hd = m->acc_hdisk[j - 1];
    double gset = 0.0;
    int iset = 0, found = 0;
    size_t c = 0;
    for (int tries = 0; tries < ORC_ACC_TRIES; ++tries) {
        double x, y, r2 = 2.0;
        while (r2 > 1.0) {
            x = ran(&R->rng); x = x + x - 1.0;
            y = ran(&R->rng); y = y + y - 1.0;
            r2 = x * x + y * y;
        }
        const double r = sqrt(r2), sp = y / r, cp = x / r;
        double g;
        if (!iset) {
            double v1, v2, rsq;
            for (;;) {
                v1 = 2.0 * ran(&R->rng) - 1.0;
                v2 = 2.0 * ran(&R->rng) - 1.0;
                rsq = v1 * v1 + v2 * v2;
                if (rsq < 1.0 && rsq > 0.0) break;
            }
            const double fac = sqrt(-2.0 * log(rsq) / rsq);
            gset = v1 * fac;
            g = v2 * fac;
            iset = 1;
        } else {
            g = gset;
            iset = 0;
        }
        p->xp = rp * cp;
        p->yp = rp * sp;
        p->zp = hd * g;
        p->rsq = p->xp * p->xp + p->yp * p->yp + p->zp * p->zp;
        p->rtot = sqrt(p->rsq);
        p->ir = orc_locate(m->r2arr, m->nrg, p->rsq);
        if (p->ir < 1 || p->ir > m->nrg - 1) continue;
        int it = orc_locate(m->costarr, m->ntg, p->zp / p->rtot);
        double lp = atan2(p->yp, p->xp);
        if (lp < 0.0) lp = lp + R2P;
        int ip = orc_locate(m->phiarr, m->npg, lp);
        p->it = it < 1 ? 1 : (it > m->ntg - 1 ? m->ntg - 1 : it);
        p->ip = ip < 1 ? 1 : (ip > m->npg - 1 ? m->npg - 1 : ip);
        c = CELL(m, p->ir, p->it, p->ip);
        if (m->rho[c] > 0.0) { found = 1; break; }
    }
    if (!found) return 0;
    *idust2 = select_dust(R, p);
    t->dbig[c] += 1;
    t->nabs[c] += 1;
    const int iop = m->species_opac[*idust2];
    if (orc_diffusing(m, p->ir, p->it)) diffuse(R, p);
    else reemit(R, p, R->tdust[c], iop);
    p->on_wall = 0;
    p->iflag = 0;
    p->i_orig = 2;
    opacset(m, p->nub, p->op);
    p->wave = 1.2398 / p->nub;
    set_kapd(p, m, 0);
    if (m->ipeel == 1)
        for (int pid = 1; pid <= m->npeel; ++pid) peeloff(R, p, p->sip, pid, *idust2);
    return 1;
}

/* newdisktherm_5_fuv.f90:294-455: one photon, stellar (ips=0, no spots)
 * or, for the last packets of a run, from the external field */
static int one_packet(orc_run *R, double *trace, int64_t i_run)
{
    const orc_model *m = R->m;
    orc_pkt p;
    memset(&p, 0, sizeof(p));
    p.on_wall = 0;
    if (i_run < m->acc_to) {
        int idust2 = 0;
        if (!emit_acc(R, &p, &idust2)) {
            R->t->counters[ORC_C_ERRORS] += 1.0;
            if (trace) {
                memset(trace, 0, 8 * sizeof(double));
                trace[3] = 1.0;
                trace[7] = -1.0;
            }
            return -1;
        }
        goto transport;
    }
    if (m->isrf_from >= 0 && i_run >= m->isrf_from) {
        emit_outside(R, &p);
        goto transport;
    }
    /* initpspot.f90:27-160; a hot-spot photon (newdisktherm_5_fuv.f90:296-313,
     * ispotph = 1) redraws its point until it lies in a spot */
    const int hot = m->ispot && i_run < m->hot_to;
    int ispotph = hot;
    p.sip = 1.0; p.sqp = 0.0; p.sup = 0.0; p.svp = 0.0;
    for (;;) {
        for (;;) {
            p.cosb = 1.0 - 2.0 * ran(&R->rng);
            p.sinb = sqrt(1.0 - p.cosb * p.cosb);
            if (p.cosb != 0.0 && fabs(p.cosb) < 1.0) break;
        }
        p.lp = R2P * ran(&R->rng);
        if (!m->ispot) break;
        const double cphi = cos(p.lp), sphi = sin(p.lp), sthet = p.sinb, cthet = p.cosb;
        const double dotsp1 = m->scspot1 * sthet * cphi + m->ssspot1 * sthet * sphi + m->cspot1 * cthet;
        const double dotsp2 = m->scspot2 * sthet * cphi + m->ssspot2 * sthet * sphi + m->cspot2 * cthet;
        if (m->nspot == 1) {
            if (ispotph == 1 && dotsp1 < m->csprad) continue;
        } else {
            if (ispotph == 1 && dotsp1 < m->csprad && dotsp2 < m->csprad) continue;
        }
        if (m->fspot < 1.0) {
            double dellon = (p.lp < PI) ? p.lp : R2P - p.lp;
            double costhp = m->cspot1 * p.cosb + m->sspot1 * p.sinb * cos(dellon);
            if (costhp > m->csprad) ispotph = 1;
            if (m->nspot == 2) {
                dellon = (p.lp < PI) ? PI - p.lp : p.lp - PI;
                costhp = m->cspot2 * p.cosb + m->sspot2 * p.sinb * cos(dellon);
                if (costhp > m->csprad) ispotph = 1;
            }
        }
        break;
    }
    initp_sample(R, &p);
    initp_transform(&p);
    if (m->ispot && (m->fspot == 1.0 || ispotph == 1)) p.nub = bbfreq(&R->rng, m->tshock);
    else p.nub = bbfreq(&R->rng, m->tstar);
    if (hot && ran(&R->rng) < 0.5)          /* half the hot-spot photons are X-rays */
        p.nub = 1.2398 / 0.06 + ran(&R->rng) * (1.2398 / 0.015 - 1.2398 / 0.06);
    p.i_orig = 1;
    opacset(m, p.nub, p.op);
    p.wave = 1.2398 / p.nub;
    set_kapd(&p, m, 0);
    double coslp = cos(p.lp), sinlp = sin(p.lp);
    p.zp = p.cosb * m->rmin;
    double rp = p.sinb * m->rmin;
    p.xp = rp * coslp;
    p.yp = rp * sinlp;
    p.rsq = p.zp * p.zp + rp * rp;
    p.rtot = sqrt(p.rsq);
    p.ux = p.sint * p.cosp; p.uy = p.sint * p.sinp; p.uz = p.cost;
    p.iflag = 0;
    p.ir = 1;
    p.on_wall = 1;
    p.it = orc_locate(m->thetarr, m->ntg, acos(p.cosb));
    p.ip = orc_locate(m->phiarr, m->npg, p.lp);
    if (m->ipeel == 1) {
        for (int pid = 1; pid <= m->npeel; ++pid) {
            int k = pid - 1;
            double cosnorm = p.cosb * m->coste[k] + (p.sinb * m->sinte[k] * (m->cospe[k] * coslp + m->sinpe[k] * sinlp));
            double sipnew = 4.0 * p.sip * cosnorm;
            if (cosnorm > 0.0) peeloff(R, &p, sipnew, pid, 0);
        }
    }
transport:;
    int nint = 0;
    int code = propagate(R, &p, &nint);
    if (code == 1) R->t->counters[ORC_C_ESCAPED] += 1.0;
    else if (code == 2) R->t->counters[ORC_C_KILLED] += 1.0;
    else R->t->counters[ORC_C_ERRORS] += 1.0;
    R->t->counters[ORC_C_HIT] += p.hit;
    R->t->counters[ORC_C_HTOT] += p.htot;
    if (trace) {
        trace[0] = p.nub; trace[1] = p.cost; trace[2] = p.phi; trace[3] = p.sip;
        trace[4] = p.sqp; trace[5] = p.sup; trace[6] = (double)nint; trace[7] = (double)code;
    }
    return code;
}

static size_t ncells(const orc_model *m) { return (size_t)(m->nrg - 1) * (m->ntg - 1) * (m->npg - 1); }
static size_t nsed(const orc_model *m) { return (size_t)10 * m->nfreq * m->nmu * m->nph * m->nap * ORC_NO; }
static size_t npsed(const orc_model *m) { return (size_t)9 * m->nfreq * m->npeel * m->nap * ORC_NO; }
static size_t nimage(const orc_model *m) { return (size_t)8 * m->nband * m->nx * m->nx * m->npeel; }

int32_t orc_run_packets(const orc_model *m, const double *tdust, orc_tallies *t,
                        uint64_t first_packet, int64_t npackets, uint64_t seed,
                        int32_t rng_mode, int32_t nthreads, double *trace)
{
    zetainit();
    if (rng_mode == 1 || nthreads <= 1) {
        orc_run R = { m, tdust, t, { rng_mode, seed, 0, 0 } };
        for (int64_t i = 0; i < npackets; ++i) {
            R.rng.packet = first_packet + (uint64_t)i;
            R.rng.n = 0;
            one_packet(&R, trace ? trace + 8 * i : NULL, i);
        }
        return 0;
    }
#ifdef _OPENMP
    /* one private tally set per thread, merged in thread order: the
     * reference's per-rank tallies followed by MPI_ALLREDUCE(SUM) */
    size_t nc = ncells(m), ns = nsed(m), nps = npsed(m), ni = nimage(m);
    orc_tallies *pt = (orc_tallies *)calloc((size_t)nthreads, sizeof(orc_tallies));
    for (int k = 0; k < nthreads; ++k) {
        pt[k].dtauabs = (double *)calloc(nc, sizeof(double));
        pt[k].nabs = (int32_t *)calloc(nc, sizeof(int32_t));
        pt[k].dbig = (int32_t *)calloc(nc, sizeof(int32_t));
        pt[k].killed = (int32_t *)calloc(nc, sizeof(int32_t));
        pt[k].sed = (float *)calloc(ns, sizeof(float));
        pt[k].psed = (float *)calloc(nps ? nps : 1, sizeof(float));
        pt[k].image = (float *)calloc(ni ? ni : 1, sizeof(float));
    }
    #pragma omp parallel num_threads(nthreads)
    {
        int k = omp_get_thread_num();
        orc_run R = { m, tdust, &pt[k], { 0, seed, 0, 0 } };
        #pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < npackets; ++i) {
            R.rng.packet = first_packet + (uint64_t)i;
            R.rng.n = 0;
            one_packet(&R, trace ? trace + 8 * i : NULL, i);
        }
    }
    for (int k = 0; k < nthreads; ++k) {
        for (size_t i = 0; i < nc; ++i) { t->dtauabs[i] += pt[k].dtauabs[i]; t->nabs[i] += pt[k].nabs[i]; t->dbig[i] += pt[k].dbig[i]; t->killed[i] += pt[k].killed[i]; }
        for (size_t i = 0; i < ns; ++i) t->sed[i] += pt[k].sed[i];
        for (size_t i = 0; i < nps; ++i) t->psed[i] += pt[k].psed[i];
        for (size_t i = 0; i < ni; ++i) t->image[i] += pt[k].image[i];
        for (int i = 0; i < 16; ++i) t->counters[i] += pt[k].counters[i];
        free(pt[k].dtauabs); free(pt[k].nabs); free(pt[k].dbig); free(pt[k].killed);
        free(pt[k].sed); free(pt[k].psed); free(pt[k].image);
    }
    free(pt);
    return 0;
#else
    return -2;
#endif
}

int32_t orc_trace_ray(const orc_model *m, double pos[3], const double u[3], int32_t ijk[3], int on_wall,
                      double tau_stop, int32_t max_steps, int32_t *faces, int32_t *cells,
                      double *tau_out, int32_t *status, uint64_t stream)
{
    /* deterministic walk: kapd from the first opacity entry at its current
     * table values (kapd = 1 per unit density when rho is an optical-depth
     * density); no RNG use, no tallies */
    orc_tallies dummy;
    memset(&dummy, 0, sizeof(dummy));
    orc_run R = { m, NULL, &dummy, { 0, 0, stream, 0 } };
    orc_pkt p;
    memset(&p, 0, sizeof(p));
    p.xp = pos[0]; p.yp = pos[1]; p.zp = pos[2];
    p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
    p.rtot = sqrt(p.rsq);
    p.ux = u[0]; p.uy = u[1]; p.uz = u[2];
    p.ir = ijk[0]; p.it = ijk[1]; p.ip = ijk[2]; p.on_wall = on_wall;
    for (int id = 0; id < 8; ++id) p.kapd[id] = 1.0;
    double tsum = 0.0, dt;
    int32_t iface, n = 0;
    *status = 3;
    while (n < max_steps) {
        testgrid(&R, &p);
        if (next_face(m, &p, p.ux, p.uy, p.uz, &dt, &iface) != 0) { *status = -1; break; }
        size_t c = CELL(m, p.ir, p.it, p.ip);
        int iop;
        double opac = cell_opac(m, &p, c, &iop);
        double dtau = opac * dt;
        if ((tsum + dtau) > tau_stop) {
            double ds = (tau_stop - tsum) / opac;
            p.xp = p.xp + ds * p.ux; p.yp = p.yp + ds * p.uy; p.zp = p.zp + ds * p.uz;
            tsum = tau_stop;
            *status = 0;
            break;
        }
        cross_face(m, &p, iface);
        p.xp = p.xp + dt * p.ux; p.yp = p.yp + dt * p.uy; p.zp = p.zp + dt * p.uz;
        p.on_wall = 1;
        p.rsq = p.xp * p.xp + p.yp * p.yp + p.zp * p.zp;
        p.rtot = sqrt(p.rsq);
        tsum = tsum + dtau;
        faces[n] = iface;
        cells[3 * n + 0] = p.ir; cells[3 * n + 1] = p.it; cells[3 * n + 2] = p.ip;
        n++;
        if (p.ir == m->nrg) { *status = 1; break; }
        if (p.ir == 0) { *status = 2; break; }
    }
    pos[0] = p.xp; pos[1] = p.yp; pos[2] = p.zp;
    ijk[0] = p.ir; ijk[1] = p.it; ijk[2] = p.ip;
    *tau_out = tsum;
    return n;
}

/* ======================================================================
 * setemiss: Planck-mean emissivity tables (dust_5.f90:507-699)
 * ====================================================================== */

/* mathsubs.f90:79-113: exp(x)-1 with the Abramowitz & Stegun 4.2.45
 * polynomial for |x| <= 0.6931472 */
double orc_expm1_as(double x)
{
    static const double c[7] = {0.9999999995, 0.4999999206, 0.1666653019, 0.0416573475,
                                0.0083013598, 0.0013298820, 0.0001413161};
    if (fabs(x) > 0.6931472) return exp(x) - 1.0;
    double s = 0.0;
    for (int i = 6; i >= 0; --i) s = x * (c[i] + s);
    return s;
}

/* mathsubs.f90:4-22 */
double orc_bnu(double nu, double T)
{
    const double x = nu / T;
    if (x == 0.0) return 0.0;
    return 0.1539897 * x * x * x / orc_expm1_as(x);
}

/* mathsubs.f90:29-47 */
double orc_dbnu(double nu, double T)
{
    const double x = nu / T;
    if (x == 0.0) return 0.0;
    return -0.1539897 * x * x * x * x / (4.0 * orc_expm1_as(x) * orc_expm1_as(-x));
}

/* module kappa_B_int (dust_5.f90:1-7) for one interval */
typedef struct { int kind; double k0, k1, nu0, dnu, teff; } orc_seg;

/* kappaB (dust_5.f90:782-794) for kind 0, chidB (:814-826) for kind 1 */
static double seg_f(const orc_seg *s, double f)
{
    const double nu = s->nu0 + f * s->dnu;
    if (s->kind == 0) return ((1.0 - f) * s->k0 + f * s->k1) * orc_bnu(nu, s->teff);
    return orc_dbnu(nu, s->teff) / ((1.0 - f) * s->k0 + f * s->k1);
}

/* trapzd (nr_2.f:233-254): the n-th refinement of the trapezoid sum in *s */
static void seg_trapzd(const orc_seg *sg, double a, double b, double *s, int n)
{
    if (n == 1) {
        *s = 0.5 * (b - a) * (seg_f(sg, a) + seg_f(sg, b));
        return;
    }
    const int it = 1 << (n - 2);
    const double tnm = (double)it;
    const double del = (b - a) / tnm;
    double x = a + 0.5 * del, sum = 0.0;
    for (int j = 1; j <= it; ++j) {
        sum = sum + seg_f(sg, x);
        x = x + del;
    }
    *s = 0.5 * (*s + (b - a) * sum / tnm);
}

/* qsimp (nr_2.f:209-231), EPS = 1e-5, JMAX = 20, over [0,1].  Returns the
 * reference's `converg` flag, which is set when the loop did NOT converge. */
int32_t orc_qsimp_seg(int32_t kind, double k0, double k1, double nu0, double dnu, double teff, double *out)
{
    const orc_seg sg = {kind, k0, k1, nu0, dnu, teff};
    double ost = -1.0e30, os = -1.0e30, st = 0.0, s = 0.0;
    for (int j = 1; j <= 20; ++j) {
        seg_trapzd(&sg, 0.0, 1.0, &st, j);
        s = (4.0 * st - ost) / 3.0;
        if (fabs(s - os) < 1.0e-5 * fabs(os)) { *out = s; return 0; }
        os = s;
        ost = st;
    }
    *out = s;
    return 1;
}

/* opacset's kappa and albedo of one species at nu (dust_5.f90:21-55); nud
 * descending, nl entries */
static void opac_ka(const double *nud, const double *kap, const double *alb, int nl, double nu,
                    double *kappa, double *albedo)
{
    if (nu >= nud[0]) { *kappa = kap[0]; *albedo = alb[0]; return; }
    if (nu <= nud[nl - 1]) { *kappa = kap[nl - 1]; *albedo = alb[nl - 1]; return; }
    int i0 = 1, i1 = nl;                                   /* 1-based, as the reference */
    while (i1 - i0 > 1) {
        const int i = (i1 + i0) / 2;
        if (nu <= nud[i - 1]) i0 = i; else i1 = i;
    }
    const double f = (nu - nud[i1 - 1]) / (nud[i0 - 1] - nud[i1 - 1]);
    const double fm1 = 1.0 - f;
    *kappa = fm1 * kap[i1 - 1] + f * kap[i0 - 1];
    *albedo = fm1 * alb[i1 - 1] + f * alb[i0 - 1];
}

/* setemiss (dust_5.f90:507-699) for every species; tables [id*nlam + k]
 * (nlambda[id] of them used, all nlam when nlambda is NULL); outputs
 * kapint[(id*nnu + j)*nT + i] = kapint(i,j,id), kappap/kappar[id*nT + i].
 * smooth[id] != 0 applies the two 21-point passes the reference gives the
 * species above ndg (:612-653).  Returns the number of qsimp calls that did
 * not converge (the reference's counta + countc), or -1 when a smoothed
 * species has fewer than 21 temperatures. */
int64_t orc_setemiss(int32_t nopac, int32_t nlam, const int32_t *nlambda, const double *nudust,
                     const double *kappad, const double *adust, int32_t nT, const double *Tint,
                     int32_t nnu, const double *nuint, const int32_t *smooth,
                     double *kapint, double *kappap, double *kappar)
{
    int64_t nfail = 0;
    if (smooth)
        for (int id = 0; id < nopac; ++id)
            if (smooth[id] && nT < 21) return -1;    /* the smoothing windows need 21 temperatures */
    double *tmp = (double *)malloc(sizeof(double) * (size_t)nT);
    for (int id = 0; id < nopac; ++id) {
        const int nl = nlambda ? nlambda[id] : nlam;
        const double *nud = nudust + (size_t)id * nlam;
        const double *kap = kappad + (size_t)id * nlam;
        const double *alb = adust + (size_t)id * nlam;
        double *K = kapint + (size_t)id * nnu * nT;
        for (int i = 0; i < nT; ++i) {
            const double teff = Tint[i];
            double nu1 = 0.0, kappa, albedo;
            opac_ka(nud, kap, alb, nl, nu1, &kappa, &albedo);
            double chi1 = kappa, kap1 = (1.0 - albedo) * kappa, kr = 0.0;
            for (int j = 0; j < nnu; ++j) {
                const double nu0 = nu1;
                nu1 = nuint[j] * teff;
                const double dnu = nu1 - nu0;
                const double dx = dnu / teff;
                const double chi0 = chi1, kap0 = kap1;
                opac_ka(nud, kap, alb, nl, nu1, &kappa, &albedo);
                chi1 = kappa;
                kap1 = (1.0 - albedo) * kappa;
                double sk = 0.0, sc = 0.0;
                if (nu0 / teff < 50.0) {
                    nfail += orc_qsimp_seg(0, kap0, kap1, nu0, dnu, teff, &sk);
                    nfail += orc_qsimp_seg(1, chi0, chi1, nu0, dnu, teff, &sc);
                }
                double *k = K + (size_t)j * nT + i;
                if (j == 0) {
                    *k = dx * sk;
                    kr = dx * sc;
                } else {
                    *k = dx * sk + K[(size_t)(j - 1) * nT + i];
                    kr = dx * sc + kr;
                }
            }
            kappar[(size_t)id * nT + i] = 1.0 / kr;
            const double tot = K[(size_t)(nnu - 1) * nT + i];
            kappap[(size_t)id * nT + i] = tot;
            for (int j = 0; j < nnu - 1; ++j) K[(size_t)j * nT + i] = K[(size_t)j * nT + i] / tot;
            K[(size_t)(nnu - 1) * nT + i] = 1.0;
        }
        if (smooth && smooth[id]) {
            double *kp = kappap + (size_t)id * nT;
            const int ns = 10;
            for (int pass = 0; pass < 2; ++pass) {
                for (int i = 1; i <= nT; ++i) {          /* 1-based i, as the reference */
                    int lo, hi;
                    if (i > ns && i <= nT - ns) { lo = i - ns; hi = i + ns; }
                    else if (i <= ns) { lo = 1; hi = 2 * i + 1; }
                    else { lo = 2 * i - nT; hi = nT; }
                    double acc = 0.0;
                    for (int ii = lo; ii <= hi; ++ii) acc = acc + kp[ii - 1];
                    tmp[i - 1] = acc / (double)(hi - lo + 1);
                }
                memcpy(kp, tmp, sizeof(double) * (size_t)nT);
            }
        }
    }
    free(tmp);
    return nfail;
}

/* ======================================================================
 * Model building: gridset's cell loop with the analytic Ulrich envelope
 * ====================================================================== */

/* zerod.f90:3-69: the root in (0,1] of xmu0^3 + (z-1) xmu0 - z xmu by the
 * trigonometric / Cardano forms.  Returns 0 when *xmu0 was set, 1 in the
 * reference's "2 roots" case and 2 when no root is positive (both leave
 * *xmu0 as it was). */
int32_t orc_zerod(double xmu, double z, double *xmu0)
{
    const double pi = 3.1415926535897932384626433;
    const double a1 = 0.0, a2 = z - 1.0, a3 = -z * xmu;
    const double q = (a1 * a1 - 3.0 * a2) / 9.0;
    const double r = (2.0 * (a1 * a1 * a1) - 9.0 * a1 * a2 + 27.0 * a3) / 54.0;
    const double q3 = q * q * q;
    if (q3 - r * r >= 0.0) {
        const double qh = sqrt(q);
        const double denom = sqrt(q3);
        const double theta = acos(r / denom);
        const double x1 = -2.0 * qh * cos(theta / 3.0) - a1 / 3.0;
        const double x2 = -2.0 * qh * cos((theta + 2.0 * pi) / 3.0) - a1 / 3.0;
        const double x3 = -2.0 * qh * cos((theta + 4.0 * pi) / 3.0) - a1 / 3.0;
        if ((x1 > 0.0 && x2 > 0.0) || (x1 > 0.0 && x3 > 0.0) || (x2 > 0.0 && x3 > 0.0)) return 1;
        int set = 0;
        if (x1 > 0.0) { *xmu0 = x1; set = 1; }
        if (x2 > 0.0) { *xmu0 = x2; set = 1; }
        if (x3 > 0.0) { *xmu0 = x3; set = 1; }
        return set ? 0 : 2;
    }
    const double factor = sqrt(r * r - q3) + fabs(r);
    const double factor3 = pow(factor, 1.0 / 3.0);
    *xmu0 = -1.0 * copysign(1.0, r) * (factor3 + q / factor3) - a1 / 3.0;
    return 0;
}

/* densenv.f90:30-45, 69-73, 111-113 without holes, gaps or bubbles
 * (ihole = igape = ibub = 0): the Ulrich (1976) infall density at radius
 * rad (R*) and polar cos/sin of a cell centre, rc the centrifugal radius,
 * rhoe0 the density scale, zero inside rmine and beyond rmax, floored at
 * rhoamb.  *flag = zerod's return (NaN density when it is not 0). */
double orc_densenv(double rad, double costin, double sina, double rc, double rhoe0, double rmine, double rmax,
                   double rhoamb, int32_t *flag)
{
    double dens = 0.0;
    *flag = 0;
    const double cosa = fabs(costin);
    const double zp = rad * cosa;
    (void)sina;
    const double xmu = zp / rad;
    if (rad > rmine) {
        const double rx = rad / rc;
        double xmu0 = NAN;
        *flag = orc_zerod(xmu, rx, &xmu0);
        if (*flag) xmu0 = NAN;
        const double factor = 1.0 / ((xmu / xmu0) + 2.0 * (xmu0 * xmu0) / rx);
        dens = rhoe0 * pow(rx, -1.5) * factor / sqrt(1.0 + xmu / xmu0);
    }
    if (rad > rmine && dens < rhoamb) dens = rhoamb;
    if (rad >= rmax) dens = 0.0;
    if (rad <= rmine) dens = 0.0;
    return dens;
}

/* gridset_5_model13new.f90:1108,1127,1146-1148: the volume of cell
 * (ir,it,ip) in cm^3, walls in R*, rmincgs = R* in cm */
double orc_cell_volume(double r1, double r2, double t1, double t2, double p1, double p2, double rmincgs)
{
    const double dr3 = pow(r2, 3.0) - pow(r1, 3.0);
    const double dcost = cos(t2) - cos(t1);
    const double dp = p2 - p1;
    return -(dr3 * dcost * dp / 3.0 * pow(rmincgs, 3.0));
}
