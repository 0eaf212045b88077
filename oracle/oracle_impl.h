/*
 * oracle_impl.h -- precision-generic body of the CPU oracle (TEST
 * INFRASTRUCTURE ONLY: it is the checker of libfpm_amd, never part of it).
 * Included twice by fpm_oracle.c with REAL = double / float and SFX = f64 / f32.
 *
 * Every function restates the reference path of THUHoloLab/Fused-Fourier-Ptycho
 * (paths relative to that repository):
 *   fft2          unnormalised C2C of cufftExecC2C (cuFPM_v3/cuFPM_pure_v3.cu:135-187)
 *   crop x pupil  cuFPM_v3/cuda/kernels.cuh:19-65
 *   loss stencil  cuFPM_v3/cuda/kernels.cuh:197-275 (helpers :5-15)
 *   back-proj.    cuFPM_v3/cuda/kernels.cuh:67-129, stitched sequentially in
 *                 LED order then pixel order (the reference uses FP32 atomics)
 *   RMSprop+clamp cuFPM_v3/cuda/kernels.cuh:340-411
 *   solve loop    cuFPM_v3/cuFPM_pure_v3.cu:128-207 (LR halving :204-206)
 *   loss value    benchmark/func_ware/ret_loss.m:3-4,13,22 ('anisotropic')
 * Arrays are column-major: element (i, j) of an R x C page at j*R + i.
 */

#define CAT2(a, b) a##_##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)
#define CT FN(cplx)

typedef struct { REAL re, im; } CT;

static inline CT FN(cmul)(CT a, CT b) { CT r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return r; }

/* ------------------------------------------------------------------ FFT */
/* Iterative radix-2 decimation-in-time on a contiguous line for powers of
 * two; for other lengths a self-sorting (Stockham) mixed-radix transform
 * with radices 4, 2, 3 and every other prime factor, each pass a direct
 * r-point sum: pass (r, n, s) maps
 *   y[q + s (r p + k)] = (sum_j x[q + s (p + j n/r)] W_r^(j k)) W_n^(p k).
 * Twiddles are computed in double and rounded once to REAL. */
typedef struct {
    int n, lg, pow2;
    int* rev;
    CT* w;    /* pow2: w[k] = exp(-2 pi i k / n), k <= n/2; else k < n */
    int nf;
    int fac[32];
} FN(fftplan);

static void FN(plan_init)(FN(fftplan) * p, int n) {
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    p->n = n;
    p->lg = lg;
    p->pow2 = (1 << lg) == n;
    p->nf = 0;
    if (!p->pow2) {
        int m = n;
        while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
        while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
        for (int q = 3; m > 1; q += 2)
            while (m % q == 0) { p->fac[p->nf++] = q; m /= q; }
        p->rev = NULL;
        p->w = (CT*)malloc(sizeof(CT) * n);
        for (int k = 0; k < n; ++k) {
            const double a = -2.0 * M_PI * (double)k / (double)n;
            p->w[k].re = (REAL)cos(a);
            p->w[k].im = (REAL)sin(a);
        }
        return;
    }
    p->rev = (int*)malloc(sizeof(int) * n);
    p->w = (CT*)malloc(sizeof(CT) * (n / 2 + 1));
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b) r |= ((i >> b) & 1) << (lg - 1 - b);
        p->rev[i] = r;
    }
    for (int k = 0; k < n / 2 + 1; ++k) {
        const double a = -2.0 * M_PI * (double)k / (double)n;
        p->w[k].re = (REAL)cos(a);
        p->w[k].im = (REAL)sin(a);
    }
}

static void FN(plan_free)(FN(fftplan) * p) {
    free(p->rev);
    free(p->w);
}

static void FN(fft_line_mixed)(const FN(fftplan) * p, CT* x, int inverse) {
    const int L = p->n;
    CT* buf = (CT*)malloc(sizeof(CT) * L);
    CT *src = x, *dst = buf;
    int n = L, s = 1;
    for (int f = 0; f < p->nf; ++f) {
        const int r = p->fac[f], m = n / r;
        for (int b = 0; b < L / r; ++b) {
            const int pp = b / s, q = b % s;
            CT a[r];
            for (int j = 0; j < r; ++j) a[j] = src[q + s * (pp + j * m)];
            for (int k = 0; k < r; ++k) {
                CT acc = {0, 0};
                for (int j = 0; j < r; ++j) {
                    CT w = p->w[((j * k) % r) * (L / r)];
                    if (inverse) w.im = -w.im;
                    const CT t = FN(cmul)(a[j], w);
                    acc.re += t.re;
                    acc.im += t.im;
                }
                CT w = p->w[((pp * k) % n) * (L / n)];
                if (inverse) w.im = -w.im;
                dst[q + s * (r * pp + k)] = FN(cmul)(acc, w);
            }
        }
        CT* t = src;
        src = dst;
        dst = t;
        n = m;
        s *= r;
    }
    if (src != x) memcpy(x, src, sizeof(CT) * L);
    free(buf);
}

static void FN(fft_line)(const FN(fftplan) * p, CT* x, int inverse) {
    if (!p->pow2) {
        FN(fft_line_mixed)(p, x, inverse);
        return;
    }
    const int n = p->n;
    for (int i = 0; i < n; ++i) {
        const int r = p->rev[i];
        if (r > i) { CT t = x[i]; x[i] = x[r]; x[r] = t; }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len) {
            for (int k = 0; k < half; ++k) {
                CT w = p->w[k * step];
                if (inverse) w.im = -w.im;
                const CT a = x[s + k];
                const CT b = FN(cmul)(x[s + k + half], w);
                x[s + k].re = a.re + b.re;
                x[s + k].im = a.im + b.im;
                x[s + k + half].re = a.re - b.re;
                x[s + k + half].im = a.im - b.im;
            }
        }
    }
}

/* unnormalised 2D transform of `batch` rows x cols pages, in place */
void FN(oracle_fft2)(CT* a, int rows, int cols, int batch, int inverse) {
    FN(fftplan) pr, pc;
    FN(plan_init)(&pr, rows);
    FN(plan_init)(&pc, cols);
    const long plane = (long)rows * cols;
    /* contiguous index (each column) */
#pragma omp parallel for schedule(static)
    for (long line = 0; line < (long)cols * batch; ++line) FN(fft_line)(&pr, a + line * rows, inverse);
    /* strided index (each row), 8 rows per block gathered into a buffer */
    int RB = 8; /* rows per block: a divisor of rows */
    while (RB > 1 && rows % RB != 0) RB /= 2;
#pragma omp parallel
    {
        CT* buf = (CT*)malloc(sizeof(CT) * (size_t)cols * RB);
#pragma omp for schedule(static)
        for (long blk = 0; blk < (long)batch * (rows / RB); ++blk) {
            const long pg = blk / (rows / RB);
            const int r0 = (int)(blk % (rows / RB)) * RB;
            CT* base = a + pg * plane;
            for (int c = 0; c < cols; ++c)
                for (int q = 0; q < RB; ++q) buf[(size_t)q * cols + c] = base[(size_t)c * rows + r0 + q];
            for (int q = 0; q < RB; ++q) FN(fft_line)(&pc, buf + (size_t)q * cols, inverse);
            for (int c = 0; c < cols; ++c)
                for (int q = 0; q < RB; ++q) base[(size_t)c * rows + r0 + q] = buf[(size_t)q * cols + c];
        }
        free(buf);
    }
    FN(plan_free)(&pr);
    FN(plan_free)(&pc);
}

/* ------------------------------------------------- one minibatch backward
 * Inputs: spectrum Shat = DFT(O) (natural order), pupil P, amplitudes Y
 * (float, page led0 + l), LED windows led[2*l] = kl, led[2*l+1] = kt.
 * Outputs: Gs (spectrum gradient, += into a zeroed N-side array), GP
 * (pupil gradient), loss (+=), margin (min |difference| at a sign decision).
 * Sign decisions of the loss stencil, one byte per pixel of the b pages
 * (page l pixel (i, j) at l nr nc + j nr + i): bit 0 = sign bit of
 * e(i, j+1) - e(i, j), bit 1 = sign bit of e(i+1, j) - e(i, j) (cyclic), the
 * two differences whose copysign values the stencil combines
 * (kernels.cuh:234-253: sign(e - e_left) at (i, j) is bit 0 of (i, j-1)).
 * force_signs (optional) replaces the stencil's own decisions with the
 * given bits; own_signs (optional) receives the own decisions; flips
 * (optional, +=) counts the bits in which they differ.
 */
static void FN(minibatch)(const CT* Shat, int Nr, int Nc, const CT* P, int nr, int nc, const float* Y,
                          const int32_t* led, int led0, int b, int pratio, CT* Gs, CT* GP, double* loss,
                          double* margin, const uint8_t* force_signs, uint8_t* own_signs, long* flips) {
    const long plane = (long)nr * nc;
    CT* sub = (CT*)malloc(sizeof(CT) * plane * b);
    CT* lat = (CT*)malloc(sizeof(CT) * plane * b);
    CT* u = (CT*)malloc(sizeof(CT) * plane * b);
    REAL* e = (REAL*)malloc(sizeof(REAL) * plane * b);

    /* crop x pupil, ifftshift folded into the store (kernels.cuh:38-64) */
#pragma omp parallel for schedule(static)
    for (int l = 0; l < b; ++l) {
        const uint32_t kl = (uint32_t)led[2 * (led0 + l)], kt = (uint32_t)led[2 * (led0 + l) + 1];
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nr; ++i) {
                const uint32_t sx = ((uint32_t)j + kl - 1u + (uint32_t)Nc / 2u) % (uint32_t)Nc;
                const uint32_t sy = ((uint32_t)i + kt - 1u + (uint32_t)Nr / 2u) % (uint32_t)Nr;
                const CT s = Shat[(long)sx * Nr + sy];
                sub[l * plane + (long)j * nr + i] = s;
                const long d = (long)((j + nc / 2) % nc) * nr + (i + nr / 2) % nr;
                lat[l * plane + d] = FN(cmul)(s, P[(long)j * nr + i]);
            }
    }
    FN(oracle_fft2)(lat, nr, nc, b, 1);

    /* loss stencil (kernels.cuh:197-275) */
    const REAL ratio = (REAL)((uint32_t)plane * (uint32_t)(pratio * pratio));
    const REAL gscale = (REAL)(pratio * pratio);
    double lsum = 0.0, mmin = INFINITY;
    long nflip = 0;
#pragma omp parallel for schedule(static) reduction(+ : lsum, nflip) reduction(min : mmin)
    for (int l = 0; l < b; ++l) {
        const float* y = Y + (long)(led0 + l) * plane;
        REAL* el = e + l * plane;
        const CT* z = lat + l * plane;
        const uint8_t* fs = force_signs ? force_signs + l * plane : NULL;
        for (long p = 0; p < plane; ++p) el[p] = M_SQRT(z[p].re * z[p].re + z[p].im * z[p].im) / ratio - (REAL)y[p];
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nr; ++i) {
                const REAL e0 = el[(long)j * nr + i];
                const REAL eL = el[(long)((j + nc - 1) % nc) * nr + i];
                const REAL eR = el[(long)((j + 1) % nc) * nr + i];
                const REAL eU = el[(long)j * nr + (i + nr - 1) % nr];
                const REAL eD = el[(long)j * nr + (i + 1) % nr];
                REAL sxl = M_COPYSIGN((REAL)1, e0 - eL), sxr = M_COPYSIGN((REAL)1, eR - e0);
                REAL syu = M_COPYSIGN((REAL)1, e0 - eU), syd = M_COPYSIGN((REAL)1, eD - e0);
                const uint8_t own = (uint8_t)((sxr < 0 ? 1 : 0) | (syd < 0 ? 2 : 0));
                if (own_signs) own_signs[l * plane + (long)j * nr + i] = own;
                if (fs) {
                    const uint8_t here = fs[(long)j * nr + i];
                    const uint8_t left = fs[(long)((j + nc - 1) % nc) * nr + i];
                    const uint8_t up = fs[(long)j * nr + (i + nr - 1) % nr];
                    sxr = (here & 1) ? (REAL)-1 : (REAL)1;
                    syd = (here & 2) ? (REAL)-1 : (REAL)1;
                    sxl = (left & 1) ? (REAL)-1 : (REAL)1;
                    syu = (up & 2) ? (REAL)-1 : (REAL)1;
                    nflip += ((here ^ own) & 1) + (((here ^ own) >> 1) & 1);
                }
                const REAL g = (sxl - sxr + syu - syd) * gscale;
                lsum += fabs((double)(eR - e0)) + fabs((double)(eD - e0));
                const double m1 = fabs((double)(eR - e0)), m2 = fabs((double)(eD - e0));
                if (m1 < mmin) mmin = m1;
                if (m2 < mmin) mmin = m2;
                const CT zz = z[(long)j * nr + i];
                const REAL mag = M_SQRT(zz.re * zz.re + zz.im * zz.im);
                CT o = {0, 0};
                if (mag > 0) {
                    const REAL inv = (REAL)1 / mag;
                    o.re = zz.re * inv * g;
                    o.im = zz.im * inv * g;
                }
                u[l * plane + (long)j * nr + i] = o;
            }
    }
    if (loss) *loss += lsum;
    if (margin && mmin < *margin) *margin = mmin;
    if (flips) *flips += nflip;
    FN(oracle_fft2)(u, nr, nc, b, 0);

    /* back-projection (kernels.cuh:87-128), sequential LED order */
    const REAL invN2 = (REAL)1 / (REAL)((uint32_t)Nr * (uint32_t)Nc);
    for (int l = 0; l < b; ++l) {
        const uint32_t kl = (uint32_t)led[2 * (led0 + l)], kt = (uint32_t)led[2 * (led0 + l) + 1];
#pragma omp parallel for schedule(static)
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nr; ++i) {
                const long sp = (long)((j + nc / 2) % nc) * nr + (i + nr / 2) % nr;
                const CT R = u[l * plane + sp];
                const CT p = P[(long)j * nr + i];
                const uint32_t sx = ((uint32_t)j + kl - 1u + (uint32_t)Nc / 2u) % (uint32_t)Nc;
                const uint32_t sy = ((uint32_t)i + kt - 1u + (uint32_t)Nr / 2u) % (uint32_t)Nr;
                CT* g = Gs + (long)sx * Nr + sy;
                g->re += (R.re * p.re + R.im * p.im) * invN2;
                g->im += (p.re * R.im - R.re * p.im) * invN2;
                const CT s = sub[l * plane + (long)j * nr + i];
                CT* q = GP + (long)j * nr + i;
                q->re += R.re * s.re + R.im * s.im;
                q->im += s.re * R.im - R.re * s.im;
            }
    }
    free(sub);
    free(lat);
    free(u);
    free(e);
}

/* Gradient of one minibatch (all L images) with the loss stencil's sign
 * decisions imposed (force_signs, optional; layout in FN(minibatch)): G_O =
 * IDFT(Gs), G_P; own_signs / flips as in FN(minibatch).  With the signs a
 * GPU run took, the FP64 result differs from that run only by arithmetic
 * (transforms, products, sums), not by near-tie decisions. */
void FN(oracle_gradient_signs)(const CT* O, int Nr, int Nc, const CT* P, int nr, int nc, const float* Y, int L,
                               const int32_t* led, int pratio, CT* G_O, CT* G_P, double* loss, double* margin,
                               const uint8_t* force_signs, uint8_t* own_signs, long* flips) {
    const long NN = (long)Nr * Nc, nn = (long)nr * nc;
    CT* Shat = (CT*)malloc(sizeof(CT) * NN);
    memcpy(Shat, O, sizeof(CT) * NN);
    FN(oracle_fft2)(Shat, Nr, Nc, 1, 0);
    memset(G_O, 0, sizeof(CT) * NN);
    memset(G_P, 0, sizeof(CT) * nn);
    if (loss) *loss = 0.0;
    if (margin) *margin = INFINITY;
    if (flips) *flips = 0;
    FN(minibatch)(Shat, Nr, Nc, P, nr, nc, Y, led, 0, L, pratio, G_O, G_P, loss, margin, force_signs, own_signs,
                  flips);
    FN(oracle_fft2)(G_O, Nr, Nc, 1, 1);
    free(Shat);
}

/* Gradient of one minibatch (all L images): G_O = IDFT(Gs), G_P. */
void FN(oracle_gradient)(const CT* O, int Nr, int Nc, const CT* P, int nr, int nc, const float* Y, int L,
                         const int32_t* led, int pratio, CT* G_O, CT* G_P, double* loss, double* margin) {
    FN(oracle_gradient_signs)(O, Nr, Nc, P, nr, nc, Y, L, led, pratio, G_O, G_P, loss, margin, NULL, NULL, NULL);
}

/* Loss sum|Dx e| + |Dy e| of all L images at (O, P) (ret_loss.m). */
double FN(oracle_loss)(const CT* O, int Nr, int Nc, const CT* P, int nr, int nc, const float* Y, int L,
                       const int32_t* led, int pratio) {
    const long NN = (long)Nr * Nc, nn = (long)nr * nc;
    CT* GO = (CT*)malloc(sizeof(CT) * NN);
    CT* GP = (CT*)malloc(sizeof(CT) * nn);
    double loss = 0.0;
    FN(oracle_gradient)(O, Nr, Nc, P, nr, nc, Y, L, led, pratio, GO, GP, &loss, NULL);
    free(GO);
    free(GP);
    return loss;
}

/* The solve of cuFPM_pure_v3.cu:128-207 with v3's update rule.  max_steps > 0
 * stops after that many minibatch steps (a bounded sample for timing).  lr
 * halves after every epoch with index >= halve_from (13 is the reference's
 * `ep > 12`, :204-206). */
void FN(oracle_solve)(const CT* O0, int Nr, int Nc, const CT* P0, int nr, int nc, const float* Y, int L,
                      const int32_t* led, const float* A, int pratio, int epochs, int batch, float lr_in,
                      float beta_in, float eps_in, float clamp_in, CT* O, CT* P, double* loss_per_epoch,
                      int max_steps, int halve_from) {
    const long NN = (long)Nr * Nc, nn = (long)nr * nc;
    memcpy(O, O0, sizeof(CT) * NN);
    memcpy(P, P0, sizeof(CT) * nn);
    REAL* mO = (REAL*)calloc(NN, sizeof(REAL));
    REAL* mP = (REAL*)calloc(nn, sizeof(REAL));
    CT* Shat = (CT*)malloc(sizeof(CT) * NN);
    CT* Gs = (CT*)malloc(sizeof(CT) * NN);
    CT* GP = (CT*)malloc(sizeof(CT) * nn);
    float lr = lr_in;
    const REAL beta = (REAL)beta_in, eps = (REAL)eps_in, am = (REAL)clamp_in;
    int steps_done = 0;
    for (int ep = 0; ep < epochs; ++ep) {
        double lsum = 0.0;
        const REAL lrr = (REAL)lr;
        for (int curr = 0; curr < L; curr += batch) {
            if (max_steps > 0 && steps_done >= max_steps) break;
            const int b = (curr + batch <= L) ? batch : L - curr;
            memcpy(Shat, O, sizeof(CT) * NN);
            FN(oracle_fft2)(Shat, Nr, Nc, 1, 0);
            memset(Gs, 0, sizeof(CT) * NN);
            memset(GP, 0, sizeof(CT) * nn);
            FN(minibatch)(Shat, Nr, Nc, P, nr, nc, Y, led, curr, b, pratio, Gs, GP, &lsum, NULL, NULL, NULL, NULL);
            FN(oracle_fft2)(Gs, Nr, Nc, 1, 1);
            /* RMSprop of O (kernels.cuh:364-382) */
#pragma omp parallel for schedule(static)
            for (long k = 0; k < NN; ++k) {
                const CT g = Gs[k];
                REAL mu = g.re * g.re + g.im * g.im;
                mu = beta * mO[k] + ((REAL)1 - beta) * mu;
                const REAL sq = lrr / (M_SQRT(mu) + eps);
                O[k].re -= g.re * sq;
                O[k].im -= g.im * sq;
                mO[k] = mu;
            }
            /* RMSprop of P + amplitude clamp (kernels.cuh:384-410) */
            for (long k = 0; k < nn; ++k) {
                const CT g = GP[k];
                REAL mu = g.re * g.re + g.im * g.im;
                mu = beta * mP[k] + ((REAL)1 - beta) * mu;
                const REAL sq = lrr / (M_SQRT(mu) + eps);
                CT w = P[k];
                w.re -= g.re * sq;
                w.im -= g.im * sq;
                const REAL a = (REAL)A[k];
                const REAL ang = M_ATAN2(w.im, w.re);
                REAL ab = M_SQRT(w.re * w.re + w.im * w.im);
                ab = a * M_FMAX(M_FMIN(ab, a + am), a - am);
                w.re = M_COS(ang) * ab;
                w.im = M_SIN(ang) * ab;
                P[k] = w;
                mP[k] = mu;
            }
            ++steps_done;
        }
        if (loss_per_epoch) loss_per_epoch[ep] = lsum;
        if (ep >= halve_from) lr = (float)((double)lr * 0.5);
        /* long solves name their progress when FPM_ORACLE_PROGRESS is set */
        if (getenv("FPM_ORACLE_PROGRESS")) {
            fprintf(stderr, "[oracle] solve %dx%d epoch %d/%d done\n", Nr, Nc, ep + 1, epochs);
            fflush(stderr);
        }
    }
    free(mO);
    free(mP);
    free(Shat);
    free(Gs);
    free(GP);
}

#undef CT
#undef FN
#undef CAT
#undef CAT2
