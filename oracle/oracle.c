/* Test infrastructure only: a float64 CPU restatement of the image-formation hot path of FastSAR, the checker the
 * HIP kernels are compared with. Nothing in the product links, loads or calls this file.
 *
 * Every stage evaluates its phases directly in float64 cycles (wrapped with x - rint(x) before sin/cos), with no
 * recurrences, so it is slower than the reference's kernels and more exact than any float32 kernel:
 *
 *   orc_rot_fir_k     rotate each parent row by each child's ramp, decimate along frequency
 *                     (fastsar/ffbp_cpu.cpp:147-255; the CUDA form fastsar/ffbp_cuda.py:48-187)
 *   orc_fir_p         decimate along pulses (fastsar/ffbp_cpu.cpp:257-293)
 *   orc_final_tiles   the per-tile separable product and the quadratic phase correction
 *                     (fastsar/ffbp_cpu.cpp:302-469; fastsar/ffbp2.py:782-870)
 *   orc_bp_points     exact backprojection onto points from oversampled range profiles, linear or cubic reads,
 *                     exact float64 ranges (fastsar/bp.py:99-138, with float64 profiles; tests/test_bp.py:23-35)
 *   orc_exact_tiles   ExactFormer's tiled algorithm (third-order range expansion about each tile centre, cubic or
 *                     linear reads of the cropped profiles) in float64 (fastsar/exact.py:240-328)
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp). Parity status of each function is stated in the oracle's Python
 * header (oracle/__init__.py). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const double TWO_PI = 6.283185307179586476925286766559;

static inline void cis_cyc(double cyc, double* c, double* s)
{
    cyc -= nearbyint(cyc);
    *c = cos(TWO_PI * cyc);
    *s = sin(TWO_PI * cyc);
}

/* xre, xim [Np, P, K]; c0, sl [Np, Cn, P]; taps [L]; yre, yim [Np * Cn, P, Ko].
 * y[n c, p, j] = sum_{r<L} taps[r] z[D j + r - pl],  z[k] = x[n, p, k] exp(2 pi i (c0 + (k - kc) sl)), 0 outside [0, K). */
void orc_rot_fir_k(const double* xre, const double* xim, int Np, int P, int K, const double* c0, const double* sl, int Cn,
                   const double* taps, int L, int pl, int D, int Ko, double kc, double* yre, double* yim)
{
    #pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int n = 0; n < Np; ++n)
        for (int c = 0; c < Cn; ++c) {
            for (int p = 0; p < P; ++p) {
                const double a = c0[((size_t)n * Cn + c) * P + p], b = sl[((size_t)n * Cn + c) * P + p];
                const double* xr = xre + ((size_t)n * P + p) * K;
                const double* xi = xim + ((size_t)n * P + p) * K;
                double* or_ = yre + (((size_t)n * Cn + c) * P + p) * Ko;
                double* oi = yim + (((size_t)n * Cn + c) * P + p) * Ko;
                for (int j = 0; j < Ko; ++j) {
                    double ar = 0.0, ai = 0.0;
                    for (int r = 0; r < L; ++r) {
                        const int k = D * j + r - pl;
                        if (k < 0 || k >= K) continue;
                        double cs, sn;
                        cis_cyc(a + ((double)k - kc) * b, &cs, &sn);
                        const double zr = xr[k] * cs - xi[k] * sn, zi = xr[k] * sn + xi[k] * cs;
                        ar += taps[r] * zr;
                        ai += taps[r] * zi;
                    }
                    or_[j] = ar;
                    oi[j] = ai;
                }
            }
        }
}

/* yre, yim [B, P, Ko]; taps [L]; zre, zim [B, Po, Ko].  z[b, i, j] = sum_{r<L} taps[r] y[b, D i + r - pl, j], 0 outside. */
void orc_fir_p(const double* yre, const double* yim, int B, int P, int Ko, const double* taps, int L, int pl, int D, int Po,
               double* zre, double* zim)
{
    #pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < B; ++b)
        for (int i = 0; i < Po; ++i) {
            double* zr = zre + ((size_t)b * Po + i) * Ko;
            double* zi = zim + ((size_t)b * Po + i) * Ko;
            for (int j = 0; j < Ko; ++j) { zr[j] = 0.0; zi[j] = 0.0; }
            for (int r = 0; r < L; ++r) {
                const int p = D * i + r - pl;
                if (p < 0 || p >= P) continue;
                const double* ar = yre + ((size_t)b * P + p) * Ko;
                const double* ai = yim + ((size_t)b * P + p) * Ko;
                for (int j = 0; j < Ko; ++j) { zr[j] += taps[r] * ar[j]; zi[j] += taps[r] * ai[j]; }
            }
        }
}

/* dre, dim [B, Pf, Qf]; ux, uy [B, Pf]; dlx, dly [T]; ucx, ucy, rc [B]; w0, gx, gyr [B, Pf] or NULL; ore, oim [B, T, T].
 * out[i][j] = sum_{p,q} d[p,q] wx_p(i) wy_p(j) exp(-2 pi i (ux_p dlx_i + uy_p dly_j)(a0 + a1 q)) exp(2 pi i qc_ij),
 * qc = (dx^2 + dy^2 - (ucx dx + ucy dy)^2) fc2 / (2 rc), wx = w0 + gx dx, wy = 1 + gyr dy (1 without weights). */
void orc_final_tiles(const double* dre, const double* dim, int B, int Pf, int Qf, int T, const double* ux, const double* uy,
                     const double* dlx, const double* dly, double a0, double a1, const double* ucx, const double* ucy,
                     const double* rc, double fc2, const double* w0, const double* gx, const double* gyr, double* ore, double* oim)
{
    #pragma omp parallel
    {
        double* Ar = (double*)malloc(sizeof(double) * (size_t)Qf * T);
        double* Ai = (double*)malloc(sizeof(double) * (size_t)Qf * T);
        double* Br = (double*)malloc(sizeof(double) * (size_t)Qf * T);
        double* Bi = (double*)malloc(sizeof(double) * (size_t)Qf * T);
        double* accr = (double*)malloc(sizeof(double) * (size_t)T * T);
        double* acci = (double*)malloc(sizeof(double) * (size_t)T * T);
        #pragma omp for schedule(dynamic, 1)
        for (int b = 0; b < B; ++b) {
            memset(accr, 0, sizeof(double) * (size_t)T * T);
            memset(acci, 0, sizeof(double) * (size_t)T * T);
            for (int p = 0; p < Pf; ++p) {
                const size_t bp = (size_t)b * Pf + p;
                const double gxp = ux[bp], gyp = uy[bp];
                for (int q = 0; q < Qf; ++q) {
                    const double f = a0 + a1 * q;
                    const double d_r = dre[bp * Qf + q], d_i = dim[bp * Qf + q];
                    for (int i = 0; i < T; ++i) {
                        double cs, sn;
                        cis_cyc(-(gxp * dlx[i]) * f, &cs, &sn);
                        const double lw = w0 ? (w0[bp] + gx[bp] * dlx[i]) : 1.0;
                        Ar[(size_t)q * T + i] = (d_r * cs - d_i * sn) * lw;
                        Ai[(size_t)q * T + i] = (d_r * sn + d_i * cs) * lw;
                    }
                    for (int j = 0; j < T; ++j) {
                        double cs, sn;
                        cis_cyc(-(gyp * dly[j]) * f, &cs, &sn);
                        const double lw = w0 ? (1.0 + gyr[bp] * dly[j]) : 1.0;
                        Br[(size_t)q * T + j] = cs * lw;
                        Bi[(size_t)q * T + j] = sn * lw;
                    }
                }
                for (int q = 0; q < Qf; ++q)
                    for (int i = 0; i < T; ++i) {
                        const double xr = Ar[(size_t)q * T + i], xi = Ai[(size_t)q * T + i];
                        double* pr = accr + (size_t)i * T;
                        double* pi = acci + (size_t)i * T;
                        const double* yr = Br + (size_t)q * T;
                        const double* yi = Bi + (size_t)q * T;
                        for (int j = 0; j < T; ++j) {
                            pr[j] += xr * yr[j] - xi * yi[j];
                            pi[j] += xr * yi[j] + xi * yr[j];
                        }
                    }
            }
            const double f = fc2 / (2.0 * rc[b]);
            for (int i = 0; i < T; ++i)
                for (int j = 0; j < T; ++j) {
                    const double dx = dlx[i], dy = dly[j], los = ucx[b] * dx + ucy[b] * dy;
                    double cs, sn;
                    cis_cyc((dx * dx + dy * dy - los * los) * f, &cs, &sn);
                    const double vr = accr[(size_t)i * T + j], vi = acci[(size_t)i * T + j];
                    ore[((size_t)b * T + i) * T + j] = vr * cs - vi * sn;
                    oim[((size_t)b * T + i) * T + j] = vr * sn + vi * cs;
                }
        }
        free(Ar); free(Ai); free(Br); free(Bi); free(accr); free(acci);
    }
}

/* Exact backprojection onto points: profiles [P, nfft] (float64, zero delay at nfft / 2, read circularly);
 * ant, ref: the pulses' antenna positions and reference ranges; pts [N, 3]; interp 1 = linear, 3 = cubic Lagrange.
 * out[n] += sum_p prof_p(t) exp(2 pi i kcyc dR),  dR = |x_n - a_p| - ref_p,  t = dR inv_dr + nfft / 2. */
void orc_bp_points(const double* pre, const double* pim, int P, int nfft, const double* ant, const double* ref,
                   const double* pts, long N, double inv_dr, double kcyc, int interp, double* ore, double* oim)
{
    #pragma omp parallel for schedule(dynamic, 64)
    for (long n = 0; n < N; ++n) {
        const double x = pts[3 * n], y = pts[3 * n + 1], z = pts[3 * n + 2];
        double are = 0.0, aim = 0.0;
        for (int p = 0; p < P; ++p) {
            const double* a = ant + 3 * (size_t)p;
            const double ax = x - a[0], ay = y - a[1], az = z - a[2];
            const double dR = sqrt(ax * ax + ay * ay + az * az) - ref[p];
            const double t = dR * inv_dr + (double)(nfft / 2);
            const double tf = floor(t), w = t - tf;
            long i = (long)tf % nfft;
            if (i < 0) i += nfft;
            const double* rr = pre + (size_t)p * nfft;
            const double* ri = pim + (size_t)p * nfft;
            double vr, vi;
            if (interp == 3) {
                const long im1 = (i - 1 + nfft) % nfft, i1 = (i + 1) % nfft, i2 = (i + 2) % nfft;
                const double wm1 = w - 1.0, wm2 = w - 2.0, wp1 = w + 1.0;
                const double k0 = -w * wm1 * wm2 / 6.0, k1 = wp1 * wm1 * wm2 / 2.0;
                const double k2 = -wp1 * w * wm2 / 2.0, k3 = wp1 * w * wm1 / 6.0;
                vr = k0 * rr[im1] + k1 * rr[i] + k2 * rr[i1] + k3 * rr[i2];
                vi = k0 * ri[im1] + k1 * ri[i] + k2 * ri[i1] + k3 * ri[i2];
            } else {
                const long i1 = (i + 1) % nfft;
                vr = rr[i] + w * (rr[i1] - rr[i]);
                vi = ri[i] + w * (ri[i1] - ri[i]);
            }
            double cs, sn;
            cis_cyc(kcyc * dR, &cs, &sn);
            are += vr * cs - vi * sn;
            aim += vr * sn + vi * cs;
        }
        ore[n] += are;
        oim[n] += aim;
    }
}

/* ExactFormer's tiled exact backprojection in float64: profiles cropped to W bins per pulse (bin m of pulse p is
 * range bin lo_p + m about the zero delay), tiles of TX x TY pixels with centres cen [ntx nty, 3]; per pixel the
 * third-order range expansion about the tile centre, du + (|d|^2 - du^2)(1/(2r) - du/(2r^2)), cubic or linear
 * read at t = (r - ref) inv_dr - lo + del inv_dr, phase kcyc (r - ref + del). Accumulates into out [nx, ny]. */
void orc_exact_tiles(const double* wre, const double* wim, int P, int W, const int* lo, const double* ant, const double* ref,
                     const double* cen, int ntx, int nty, int TX, int TY, int nx, int ny, double sx, double sy,
                     const double* e1, const double* e2, double inv_dr, double kcyc, int cubic, double* ore, double* oim)
{
    #pragma omp parallel for schedule(dynamic, 1)
    for (long b = 0; b < (long)ntx * nty; ++b) {
        const int bx = (int)(b / nty), by = (int)(b % nty);
        const double cx = cen[3 * b], cy = cen[3 * b + 1], cz = cen[3 * b + 2];
        for (int q = 0; q < TX * TY; ++q) {
            const int li = q / TY, lj = q % TY, ix = bx * TX + li, iy = by * TY + lj;
            if (ix >= nx || iy >= ny) continue;
            const double a = (li - 0.5 * (TX - 1)) * sx, c = (lj - 0.5 * (TY - 1)) * sy;
            const double dx = a * e1[0] + c * e2[0], dy = a * e1[1] + c * e2[1], dz = a * e1[2] + c * e2[2];
            const double dd = dx * dx + dy * dy + dz * dz;
            double are = 0.0, aim = 0.0;
            for (int p = 0; p < P; ++p) {
                const double* A = ant + 3 * (size_t)p;
                const double wx = cx - A[0], wy = cy - A[1], wz = cz - A[2];
                const double r = sqrt(wx * wx + wy * wy + wz * wz), dR = r - ref[p];
                const double ux = wx / r, uy = wy / r, uz = wz / r;
                const double du = dx * ux + dy * uy + dz * uz;
                const double del = du + (dd - du * du) * (0.5 / r - du * 0.5 / (r * r));
                const double t = (dR + del) * inv_dr - lo[p];
                const double tf = floor(t), w = t - tf;
                const long i = (long)tf;
                const double* rr = wre + (size_t)p * W;
                const double* ri = wim + (size_t)p * W;
                double vr, vi;
                if (cubic) {
                    const double wm1 = w - 1.0, wm2 = w - 2.0, wp1 = w + 1.0;
                    const double k0 = -w * wm1 * wm2 / 6.0, k1 = wp1 * wm1 * wm2 / 2.0;
                    const double k2 = -wp1 * w * wm2 / 2.0, k3 = wp1 * w * wm1 / 6.0;
                    vr = k0 * rr[i - 1] + k1 * rr[i] + k2 * rr[i + 1] + k3 * rr[i + 2];
                    vi = k0 * ri[i - 1] + k1 * ri[i] + k2 * ri[i + 1] + k3 * ri[i + 2];
                } else {
                    vr = rr[i] + w * (rr[i + 1] - rr[i]);
                    vi = ri[i] + w * (ri[i + 1] - ri[i]);
                }
                double cs, sn;
                cis_cyc(kcyc * (dR + del), &cs, &sn);
                are += vr * cs - vi * sn;
                aim += vr * sn + vi * cs;
            }
            ore[(size_t)ix * ny + iy] += are;
            oim[(size_t)ix * ny + iy] += aim;
        }
    }
}
