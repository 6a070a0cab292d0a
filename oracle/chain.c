/*
 * FP64 C restatement of the magnification-PDF hot path (checker only).
 *
 * TEST INFRASTRUCTURE: loaded (ctypes) only by tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg, as the checker.  The product path never
 * links it.
 *
 * What it restates, function by function:
 *   nfw_Fh            sgl.py:694-720 (Wright & Brainerd F, h; series and
 *                     cancellation-free branches)
 *   tc_*              sgl_full.py:740-798 (truncated-NFW correction I1, I2)
 *   trunc_kg          sgl_full.py:660-669 (kappa, gamma of NFW cut at x_t)
 *   oracle_cells      one cell per (slice, mass, impact-parameter bin):
 *                     jump xi = -ln detA (sgl.py:808-815), source-plane
 *                     Esscher tilt w e^{-xi} (sgl.py:949-951)
 *   oracle_cells_ell  the same with halo ellipticity: n_phi orientation nodes
 *                     per cell, pseudo-elliptical kappa, gamma
 *                     (sgl_full.py:582-613, 1659, 2038-2064)
 *   oracle_cells_fil  one cell per (slice, mass, u bin, orientation) of the
 *                     filament population (sgl_full.py:1984-2007)
 *   oracle_lnphi      compensated compound-Poisson exponent
 *                     Lambda(k) = sum_c w_c (e^{ik xi_c} - 1 - ik xi_c)
 *                     with the reference's cancellation-free forms
 *                     (sgl_full.py:2599-2613: cos-1 = -2 sin^2(ph/2),
 *                     sin-ph by series below |ph| = 1e-2), plus the
 *                     per-slice amplitudes a_i(k) = sum beta w (e^{ik xi}-1)
 *                     of the two-point term (sgl_full.py:3353-3397)
 *   oracle_invert     P(xi) = (1/pi) Re int_0^inf e^{-ik xi} phi(k) dk
 *                     (sgl.py:994-1012) with phi linear between k nodes
 *                     (Filon weights), phi(0) = 1
 *   oracle_sub_ring   ring average of a host's projected clump density and
 *                     its mean clump-ray cosine (subhalos.py:287-330)
 *   oracle_sub_emit   the points of decorated host cells on the rebuilt host
 *                     (subhalos_exact.py:802-925)
 *   oracle_bin_fine   Cells.add's fine (kappa, gamma) lattice sums
 *                     (sgl_full.py:1314-1383, 971-1046)
 *   oracle_xi_deposit the xi_lin deposit of one row's points on the coarse
 *                     lattice (sgl_full.py:1081-1183, 1397-1441)
 *   oracle_xi_rate    the positive closure's lens rate per coarse bin
 *                     (sgl_full.py:1270-1280)
 *   oracle_vp_lnr     ln R of the positive closure on a tensor grid: group
 *                     sums through the rates' interval form and the chain
 *                     recursion (pos_closure_fused.py:52-140,
 *                     pos_closure.py:183-240)
 *
 * Build: oracle/build.py (gcc -O2 -fopenmp -shared).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NFW_W_SERIES 0.25
#define NFW_N_SERIES 28
#define NFW_X_H 0.9
#define TC_A_SMALL 1e-3
#define TC_BAND 1e-5

static void nfw_Fh(double x, double *F, double *h) {
    double w = (1.0 - x) * (1.0 + x);
    if (fabs(w) < NFW_W_SERIES) {
        double T = 0.0;
        for (int n = NFW_N_SERIES - 1; n >= 0; --n) T = T * w + 1.0 / (2 * n + 3);
        *F = T;
        *h = log(0.5 * x) + (1.0 + w * T);
        return;
    }
    if (x < 1.0) {
        double e = (1.0 - x) / x, s = sqrt(w);
        double A = log1p(e + sqrt(e * (e + 2.0)));
        *F = (A / s - 1.0) / w;
        if (x < NFW_X_H)
            *h = (log(2.0 / x) * x * x / (1.0 + s) + log1p(-x * x / (2.0 * (1.0 + s)))) / s;
        else
            *h = log(0.5 * x) + A / s;
        return;
    }
    double v = -w, sv = sqrt(v), B = atan(sv);
    *F = (1.0 - B / sv) / v;
    *h = log(0.5 * x) + B / sv;
}

static void tc_closed_raw(double x, double c, double *I1, double *I2) {
    double q = sqrt((c - x) * (c + x));
    double d = (x - 1.0) * (x - c) / (x * (1.0 + c));
    double L = log((c + q) / (2.0 * (1.0 + c))) - (q - c) / (1.0 + c);
    if (x < 1.0) {
        double u = (1.0 - x) * (1.0 + x), e = (1.0 - x) / x;
        double A0 = log1p(e + sqrt(e * (e + 2.0)));
        double At = log1p(d + sqrt(d * (d + 2.0)));
        double su = sqrt(u);
        *I1 = (q / (1.0 + c) - 1.0) / u + (A0 - At) / (u * su);
        *I2 = (L + (A0 - At) / su) / (x * x);
        return;
    }
    double v = (x - 1.0) * (x + 1.0);
    double B0 = 2.0 * asin(sqrt((x - 1.0) / (2.0 * x)));
    double md = -d > 0.0 ? -d : 0.0;
    double Bt = 2.0 * asin(sqrt(md / 2.0));
    double sv = sqrt(v);
    *I1 = (1.0 - q / (1.0 + c)) / v - (B0 - Bt) / (v * sv);
    *I2 = (L + (B0 - Bt) / sv) / (x * x);
}

/* 24-node Gauss-Legendre on [0,1]; nodes/weights filled once from the
 * Golub-Welsch-free Newton iteration on P_24 (closed-form recurrence). */
static double GLX[24], GLW[24];
static int gl_ready = 0;
static void gl_init(void) {
    const int n = 24;
    for (int i = 0; i < n; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1.0);
            double z1 = z;
            z = z1 - p1 / pp;
            if (fabs(z - z1) < 1e-16) break;
        }
        /* map [-1,1] -> [0,1] as sgl_full.py:635 */
        GLX[n - 1 - i] = 0.5 * (z + 1.0);
        GLW[n - 1 - i] = 0.5 * 2.0 / ((1.0 - z * z) * pp * pp);
    }
    gl_ready = 1;
}

static void tc_gl(double a, double xt, double *s1o, double *s2o) {
    double s1 = 0.0, s2 = 0.0;
    for (int j = 0; j < 24; ++j) {
        double v = GLX[j], s = 1.0 - v * v;
        double d = 1.0 - a * a * s * s;
        if (d < 1e-300) d = 1e-300;
        s1 += GLW[j] * 2.0 * v * s / ((s + xt) * (s + xt) * sqrt(d));
        double e = 1.0 - a * a * v * v;
        if (e < 0.0) e = 0.0;
        s2 += GLW[j] * v / ((v + xt) * (v + xt) * (1.0 + sqrt(e)));
    }
    *s1o = s1;
    *s2o = s2;
}

static void tc_point(double x, double xt, double *I1, double *I2) {
    double a = x / xt;
    if (a >= 1.0) { *I1 = 0.0; *I2 = 0.0; return; }
    if (a < TC_A_SMALL) {
        double i0 = log((1.0 + xt) / xt) - 1.0 / (1.0 + xt), u1 = 1.0 + xt;
        double jj = 0.5 * (u1 * u1 - xt * xt) - 3.0 * xt * (u1 - xt) + 3.0 * xt * xt * log(u1 / xt)
                    + xt * xt * xt * (1.0 / u1 - 1.0 / xt);
        *I1 = i0 + 0.5 * a * a * jj;
        *I2 = 0.5 * i0 + 0.125 * a * a * jj;
        return;
    }
    if (fabs(x - 1.0) < TC_BAND) {
        if (xt <= 1.0 + 2.0 * TC_BAND) { tc_gl(a, xt, I1, I2); return; }
        double a1, b1, a2, b2;
        tc_closed_raw(1.0 - TC_BAND, xt, &a1, &b1);
        tc_closed_raw(1.0 + TC_BAND, xt, &a2, &b2);
        double f = (x - (1.0 - TC_BAND)) / (2.0 * TC_BAND);
        *I1 = a1 + f * (a2 - a1);
        *I2 = b1 + f * (b2 - b1);
        return;
    }
    tc_closed_raw(x, xt, I1, I2);
}

void oracle_trunc_kg(int64_t n, const double *x, double ks, double xt, double *kap, double *gam) {
    if (!gl_ready) gl_init();
    for (int64_t i = 0; i < n; ++i) {
        double F, h, xi = x[i];
        nfw_Fh(xi, &F, &h);
        double ku = 2.0 * F, gu = 4.0 * h / (xi * xi) - ku, k, kb;
        if (xi < xt) {
            double I1, I2;
            tc_point(xi, xt, &I1, &I2);
            k = ku - 2.0 * I1;
            kb = ku + gu - 4.0 * I2;
        } else {
            double mut = log(1.0 + xt) - xt / (1.0 + xt);
            k = 0.0;
            kb = 4.0 * mut / (xi * xi);
        }
        kap[i] = ks * k;
        gam[i] = ks * (kb - k);
    }
}

void oracle_nfw_Fh(int64_t n, const double *x, double *F, double *h) {
    for (int64_t i = 0; i < n; ++i) nfw_Fh(x[i], &F[i], &h[i]);
}

/*
 * Cells of the kernel-shape population.  rows: [n_rows][6] =
 * (k_s, r_s, x_t, r_vir, w_row, beta) with w_row = dchi (1+z)^2 dn [Mpc^-2],
 * r_s and r_vir physical Mpc; u_edges[n_b+1] in units of r_vir.
 * Outputs per cell (row-major [n_rows][n_b]): xi, ws = w e^{-xi} (0 for a
 * dropped cell), kap, gam.  A cell is dropped when detA <= 0 (sgl.py:812-815)
 * or xi > XI_MAX (the reference's jump grid ends at 30, sgl.py:744).
 */
#define XI_MAX 30.0
void oracle_cells(int64_t n_rows, const double *rows, int64_t n_b, const double *u_edges,
                  double *xi, double *ws, double *kap, double *gam) {
    if (!gl_ready) gl_init();
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < n_rows; ++r) {
        const double *R = rows + 6 * r;
        double ks = R[0], rs = R[1], xt = R[2], rvir = R[3], wrow = R[4];
        for (int64_t l = 0; l < n_b; ++l) {
            double b0 = rvir * u_edges[l], b1 = rvir * u_edges[l + 1];
            double bc = sqrt(b0 * b1), x = bc / rs;
            double area = M_PI * (b1 * b1 - b0 * b0);
            double k, g;
            oracle_trunc_kg(1, &x, ks, xt, &k, &g);
            double detA = (1.0 - k) * (1.0 - k) - g * g;
            int64_t o = r * n_b + l;
            kap[o] = k;
            gam[o] = g;
            if (!(detA > 0.0)) { xi[o] = 0.0; ws[o] = 0.0; continue; }
            double x_ = -log(detA);
            if (x_ > XI_MAX) { xi[o] = 0.0; ws[o] = 0.0; continue; }
            xi[o] = x_;
            ws[o] = wrow * area * detA;     /* Esscher tilt e^{-xi} = detA */
        }
    }
}

/*
 * Pseudo-elliptical truncated NFW (lensing.cpp:56 kappagammaNFWeps as the
 * reference transcribes it, sgl_full.py:582-613): the circular profile at the
 * elliptical radius x_e of x1 = sqrt(1-eps) cos(phi) x, x2 = sqrt(1+eps)
 * sin(phi) x, with c2 = (x1^2 - x2^2) / x_e^2,
 *   kappa   = k0 + eps c2 g0,
 *   gamma^2 = g0^2 + 2 eps c2 g0 k0 + eps^2 (k0^2 - (c2 g0)^2)
 * (the engine's cos^2 form of the eps^2 term, kept by the reference), g0 = ks
 * below x_e = 1e-4; beyond the truncation (x_e >= x_t) kappa = 0 and gamma
 * is the monopole shear 4 ks mu(x_t) / x^2 at the circular x.
 */
static void kg_eps(double eps, double ks, double x, double phi, double xt, double *k, double *g) {
    double x1 = sqrt(1.0 - eps) * cos(phi) * x, x2 = sqrt(1.0 + eps) * sin(phi) * x;
    double xe = fmax(sqrt(x1 * x1 + x2 * x2), 1e-12);
    double c2 = (x1 * x1 - x2 * x2) / fmax(xe * xe, 1e-300);
    double k0, g0;
    oracle_trunc_kg(1, &xe, ks, xt, &k0, &g0);
    if (xe < 1e-4) g0 = ks;
    double kk = k0 + eps * c2 * g0;
    double g2 = g0 * g0 + 2.0 * eps * c2 * g0 * k0 + eps * eps * (k0 * k0 - (c2 * g0) * (c2 * g0));
    if (xe >= xt) {
        double gpm = ks * 4.0 * (log(1.0 + xt) - xt / (1.0 + xt)) / fmax(x * x, 1e-300);
        kk = 0.0;
        g2 = gpm * gpm;
    }
    *k = kk;
    *g = sqrt(fmax(g2, 0.0));
}

void oracle_kg_eps(int64_t n, const double *x, const double *phi, double eps, double ks, double xt,
                   double *kap, double *gam) {
    if (!gl_ready) gl_init();
    for (int64_t i = 0; i < n; ++i) kg_eps(eps, ks, x[i], phi[i], xt, &kap[i], &gam[i]);
}

/*
 * Cells with halo ellipticity (the "ell" ingredient): every impact-parameter
 * cell of a row splits into n_phi orientation nodes phi_q = (q + 1/2)
 * (pi/2) / n_phi, each with 1/n_phi of the cell's weight and the
 * pseudo-elliptical kappa, gamma of the row's eps (sgl_full.py:1659, 2038-2064).
 * Output order [n_rows][n_b][n_phi]; jump rule and tilt as oracle_cells.
 */
void oracle_cells_ell(int64_t n_rows, const double *rows, const double *eps, int64_t n_b,
                      const double *u_edges, int64_t n_phi, double *xi, double *ws, double *kap,
                      double *gam) {
    if (!gl_ready) gl_init();
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < n_rows; ++r) {
        const double *R = rows + 6 * r;
        double ks = R[0], rs = R[1], xt = R[2], rvir = R[3], wrow = R[4];
        for (int64_t l = 0; l < n_b; ++l) {
            double b0 = rvir * u_edges[l], b1 = rvir * u_edges[l + 1];
            double bc = sqrt(b0 * b1), x = bc / rs;
            double area = M_PI * (b1 * b1 - b0 * b0) / (double)n_phi;
            for (int64_t q = 0; q < n_phi; ++q) {
                double phi = (q + 0.5) * (0.5 * M_PI / (double)n_phi), k, g;
                kg_eps(eps[r], ks, x, phi, xt, &k, &g);
                double detA = (1.0 - k) * (1.0 - k) - g * g;
                int64_t o = (r * n_b + l) * n_phi + q;
                kap[o] = k;
                gam[o] = g;
                if (!(detA > 0.0)) { xi[o] = 0.0; ws[o] = 0.0; continue; }
                double x_ = -log(detA);
                if (x_ > XI_MAX) { xi[o] = 0.0; ws[o] = 0.0; continue; }
                xi[o] = x_;
                ws[o] = wrow * area * detA;
            }
        }
    }
}

/*
 * Filament cells (the filament block of the reference's population builder,
 * sgl_full.py:1984-2007, with the cylinder profile of filaments.py:54-66):
 * per row (k0base, rs, w_row, beta) -- kappa amplitude k0 = k0base /
 * max(2 rs, |cos phi| L), L = 20 rs -- one cell per (u bin, orientation node):
 * u bins from u_edges[n_u + 1] (first edge 0, centre u/2 there, else the
 * geometric centre), orientations phi[n_phi] with weights wphi[n_phi];
 * kappa = 2 pi k0 / s (u <= 1), gamma = 4 pi k0 / (s + 1) - kappa,
 * s = sqrt(1 + u^2); weight w_row pi rs^2 (u1^2 - u0^2) wphi; same jump rule
 * and tilt as the halo cells.  Output cell order: row, u bin, orientation.
 */
void oracle_cells_fil(int64_t n_rows, const double *rows, int64_t n_u, const double *u_edges,
                      int64_t n_phi, const double *phi, const double *wphi, double *xi,
                      double *ws, double *kap, double *gam) {
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < n_rows; ++r) {
        const double *R = rows + 4 * r;
        double k0base = R[0], rs = R[1], wrow = R[2];
        for (int64_t l = 0; l < n_u; ++l) {
            double u0 = u_edges[l], u1 = u_edges[l + 1];
            double uc = (u0 > 0.0) ? sqrt(u0 * u1) : 0.5 * u1;
            double s = sqrt(1.0 + uc * uc);
            double area = M_PI * rs * rs * (u1 * u1 - u0 * u0);
            for (int64_t n = 0; n < n_phi; ++n) {
                double Lf = 20.0 * rs, c = fabs(cos(phi[n])) * Lf;
                double k0 = k0base / (2.0 * rs > c ? 2.0 * rs : c);
                double k = (uc <= 1.0) ? 2.0 * M_PI * k0 / s : 0.0;
                double g = 4.0 * M_PI * k0 / (s + 1.0) - k;
                int64_t o = (r * n_u + l) * n_phi + n;
                kap[o] = k;
                gam[o] = g;
                double detA = (1.0 - k) * (1.0 - k) - g * g;
                if (!(detA > 0.0)) { xi[o] = 0.0; ws[o] = 0.0; continue; }
                double x_ = -log(detA);
                if (x_ > XI_MAX) { xi[o] = 0.0; ws[o] = 0.0; continue; }
                xi[o] = x_;
                ws[o] = wrow * area * wphi[n] * detA;
            }
        }
    }
}

/*
 * lam[2*n_k] (re, im) = sum_c ws_c (e^{ik xi_c} - 1 - ik xi_c);
 * amp[n_slices][2*n_k] = sum_{c in slice} beta_c ws_c (e^{ik xi_c} - 1).
 * Cells are given per slice: slice_off[n_slices+1] offsets into xi/ws/bws.
 */
void oracle_lnphi(int64_t n_slices, const int64_t *slice_off, const double *xi, const double *ws,
                  const double *bws, int64_t n_k, const double *k, double *lam, double *amp) {
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t n = 0; n < n_k; ++n) {
        double kn = k[n], lr = 0.0, li = 0.0;
        for (int64_t s = 0; s < n_slices; ++s) {
            double ar = 0.0, ai = 0.0;
            for (int64_t c = slice_off[s]; c < slice_off[s + 1]; ++c) {
                double w = ws[c];
                if (w == 0.0) continue;
                double ph = kn * xi[c];
                double sn = sin(ph), hs = sin(0.5 * ph);
                double cm1 = -2.0 * hs * hs;
                double smp;
                if (fabs(ph) < 1e-2) {
                    double p2 = ph * ph;
                    smp = -ph * p2 / 6.0 * (1.0 - p2 / 20.0 * (1.0 - p2 / 42.0));
                } else {
                    smp = sn - ph;
                }
                lr += w * cm1;
                li += w * smp;
                ar += bws[c] * cm1;
                ai += bws[c] * sn;
            }
            amp[s * 2 * n_k + 2 * n] = ar;
            amp[s * 2 * n_k + 2 * n + 1] = ai;
        }
        lam[2 * n] = lr;
        lam[2 * n + 1] = li;
    }
}

/* E1(th) = int_0^1 s e^{-i th s} ds, E0(th) = int_0^1 e^{-i th s} ds. */
static void filon_e(double th, double *e0r, double *e0i, double *e1r, double *e1i) {
    if (fabs(th) < 0.5) {
        /* sum_m (-i th)^m / (m! (m+1)), (m! (m+2)) */
        double tr = 1.0, ti = 0.0; /* (-i th)^m / m! */
        double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
        for (int m = 0; m < 30; ++m) {
            a0r += tr / (m + 1);
            a0i += ti / (m + 1);
            a1r += tr / (m + 2);
            a1i += ti / (m + 2);
            /* multiply by (-i th)/(m+1) */
            double nr = ti * th / (m + 1), ni = -tr * th / (m + 1);
            tr = nr;
            ti = ni;
        }
        *e0r = a0r; *e0i = a0i; *e1r = a1r; *e1i = a1i;
        return;
    }
    double c = cos(th), s = sin(th);
    /* E0 = (1 - e^{-i th}) / (i th) = (sin th + i (cos th - 1)) / th */
    *e0r = s / th;
    *e0i = (c - 1.0) / th;
    /* E1 = e^{-i th}(i/th + 1/th^2) - 1/th^2 */
    double ar = 1.0 / (th * th), ai = 1.0 / th;
    *e1r = c * ar + s * ai - 1.0 / (th * th);
    *e1i = c * ai - s * ar;
}

/*
 * Filon weights: A_n(t) = int hat_n(k) e^{-ikt} dk on nodes kq[0..nq-1]
 * (kq[0] = 0), hat_n the piecewise-linear basis, the last node's hat
 * one-sided.  wts[j][2*nq]: (Re A_n / pi, -Im A_n / pi), so that
 * p(t_j) = sum_n wts[j][2n] Re phi_n + wts[j][2n+1] Im phi_n.
 */
void oracle_filon_weights(int64_t n_t, const double *t, int64_t nq, const double *kq, double *wts) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n_t; ++j) {
        double tj = t[j];
        double *W = wts + j * 2 * nq;
        for (int64_t n = 0; n < 2 * nq; ++n) W[n] = 0.0;
        for (int64_t n = 0; n + 1 < nq; ++n) {
            double a = kq[n], h = kq[n + 1] - a, th = h * tj;
            double e0r, e0i, e1r, e1i;
            filon_e(th, &e0r, &e0i, &e1r, &e1i);
            double cr = cos(a * tj), ci = -sin(a * tj);   /* e^{-i a t} */
            /* up half (node n+1): h e^{-iat} E1 ; down half (node n): h e^{-iat} (E0 - E1) */
            double ur = h * (cr * e1r - ci * e1i), ui = h * (cr * e1i + ci * e1r);
            double dr_ = e0r - e1r, di_ = e0i - e1i;
            double dr = h * (cr * dr_ - ci * di_), di = h * (cr * di_ + ci * dr_);
            W[2 * n] += dr / M_PI;
            W[2 * n + 1] += -di / M_PI;
            W[2 * (n + 1)] += ur / M_PI;
            W[2 * (n + 1) + 1] += -ui / M_PI;
        }
    }
}

/* p[j] = sum_n wts(t_j)[2n] Re phi_n + wts[2n+1] Im phi_n, weights on the fly. */
void oracle_invert(int64_t n_t, const double *t, int64_t nq, const double *kq, const double *phi,
                   double *p) {
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n_t; ++j) {
        double tj = t[j], acc = 0.0;
        for (int64_t n = 0; n + 1 < nq; ++n) {
            double a = kq[n], h = kq[n + 1] - a, th = h * tj;
            double e0r, e0i, e1r, e1i;
            filon_e(th, &e0r, &e0i, &e1r, &e1i);
            double cr = cos(a * tj), ci = -sin(a * tj);
            double ur = h * (cr * e1r - ci * e1i), ui = h * (cr * e1i + ci * e1r);
            double dr_ = e0r - e1r, di_ = e0i - e1i;
            double dr = h * (cr * dr_ - ci * di_), di = h * (cr * di_ + ci * dr_);
            /* Re(phi * A) = Re phi Re A - Im phi Im A */
            acc += phi[2 * n] * dr - phi[2 * n + 1] * di;
            acc += phi[2 * (n + 1)] * ur - phi[2 * (n + 1) + 1] * ui;
        }
        p[j] = acc / M_PI;
    }
}

/* ------------------------------------------------------------------------
 * Subhalo sector (exact decorated host, moments method).
 *
 * oracle_sub_ring: the ring average of a host's projected clump density and
 * the mean cosine of the clump-ray angle it weights
 * (subhalos.py:287-330, _ring_and_pair_angle_mean_cos):
 *   ring[i][k] = (1/nt) sum_t n(R_t),  ecd[i][k] = sum_t n(R_t) cth_t / sum_t n(R_t),
 *   R_t = sqrt(r_i^2 + d_k^2 - 2 r_i d_k cth_t), n the linear interpolant of
 *   (xp, fp) on the uniform grid xp (fp[0] below, 0 above, np.interp's
 *   fused slope (x - xp_j) + fp_j).
 */
void oracle_sub_ring(int64_t nr, const double *r, int64_t nd, const double *d, int64_t nt,
                     const double *cth, int64_t n, const double *xp, const double *fp,
                     double *ring, double *ecd) {
    double x0 = xp[0], xl = xp[n - 1], inv_dx = (n - 1) / (xl - x0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nr; ++i) {
        for (int64_t k = 0; k < nd; ++k) {
            double c = r[i] * r[i] + d[k] * d[k], trd = 2.0 * r[i] * d[k];
            double s = 0.0, sc = 0.0;
            for (int64_t t = 0; t < nt; ++t) {
                double arg = c - trd * cth[t];
                double x = arg > 0.0 ? sqrt(arg) : 0.0, v;
                if (x <= x0) {
                    v = fp[0];
                } else if (x >= xl) {
                    v = (x == xl) ? fp[n - 1] : 0.0;
                } else {
                    int64_t j = (int64_t)((x - x0) * inv_dx);
                    if (j > n - 2) j = n - 2;
                    if (j < 0) j = 0;
                    double sl = (fp[j + 1] - fp[j]) / (xp[j + 1] - xp[j]);
                    v = fma(sl, x - xp[j], fp[j]);
                }
                s += v;
                sc += v * cth[t];
            }
            ring[i * nd + k] = s / (double)nt;
            ecd[i * nd + k] = s > 0.0 ? sc / s : 0.0;
        }
    }
}

/*
 * oracle_sub_emit: the points of decorated host cells (subhalos_exact.py:
 * 802-925, _moments_fused_general).  For every cell (x, phi, W) and node q:
 * the coarse-grid tables [n_g][nq] (K offset, removed mass S, weight, shear
 * cross term A1, shear square A2) interpolated linearly in ln x (clamped),
 * the host rebuilt at max(M - S, Mmin) with (ln r_s, ln k_s, eps) linear in
 * ln M on the table (clamped), kept at the host's physical truncation
 * radius, its pseudo-elliptical truncated kappa_h, gamma_h at the cell
 * (kg_eps), and
 *   K = kappa_h + K_q,  G = sqrt(max(gamma_h^2 + 2 gamma_h A1 + A2, 0)),  W_q = W w_q.
 * phq NULL: circular (phi = 0, eps = 0).  Output [n_cell][nq].
 */
void oracle_sub_emit(int64_t n_cell, const double *xq, const double *phq, const double *W,
                     double rs, double xt, int64_t nq, int64_t n_g, double xg0, const double *Kp,
                     const double *Sp, const double *wp, const double *A1, const double *A2,
                     const double *lxg, int64_t m_t, const double *lMt, const double *log_rs,
                     const double *log_ks, const double *ept, double M, double Mmin, double *Ko,
                     double *Go, double *Wo) {
    if (!gl_ready) gl_init();
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n_cell; ++i) {
        double x = xq[i], phi = phq ? phq[i] : 0.0;
        double lx = x > xg0 ? log(x) : lxg[0];
        if (lx > lxg[n_g - 1]) lx = lxg[n_g - 1];
        int64_t lo = 0, hi = n_g - 1;
        while (lo < hi - 1) {
            int64_t mid = (lo + hi) >> 1;
            if (lxg[mid] <= lx) lo = mid; else hi = mid;
        }
        double tw = (lx - lxg[lo]) / (lxg[lo + 1] - lxg[lo]), om = 1.0 - tw;
        for (int64_t q = 0; q < nq; ++q) {
            int64_t a = lo * nq + q, b = (lo + 1) * nq + q;
            double kp = om * Kp[a] + tw * Kp[b], sp = om * Sp[a] + tw * Sp[b];
            double w = om * wp[a] + tw * wp[b], a1 = om * A1[a] + tw * A1[b];
            double a2 = om * A2[a] + tw * A2[b];
            double Me = M - sp;
            if (Me < Mmin) Me = Mmin;
            double lMe = log(Me), rse, kse, epe;
            if (lMe <= lMt[0]) {
                rse = exp(log_rs[0]); kse = exp(log_ks[0]); epe = ept[0];
            } else if (lMe >= lMt[m_t - 1]) {
                rse = exp(log_rs[m_t - 1]); kse = exp(log_ks[m_t - 1]); epe = ept[m_t - 1];
            } else {
                int64_t l = 0, h = m_t - 1;
                while (l < h - 1) {
                    int64_t mid = (l + h) >> 1;
                    if (lMt[mid] <= lMe) l = mid; else h = mid;
                }
                double f = (lMe - lMt[l]) / (lMt[l + 1] - lMt[l]);
                rse = exp(log_rs[l] + f * (log_rs[l + 1] - log_rs[l]));
                kse = exp(log_ks[l] + f * (log_ks[l + 1] - log_ks[l]));
                epe = ept[l] + f * (ept[l + 1] - ept[l]);
            }
            double kh, gh;
            kg_eps(phq ? epe : 0.0, kse, x * rs / rse, phi, xt * rs / rse, &kh, &gh);
            double g2 = gh * gh + 2.0 * gh * a1 + a2;
            Ko[i * nq + q] = kh + kp;
            Go[i * nq + q] = g2 > 0.0 ? sqrt(g2) : 0.0;
            Wo[i * nq + q] = W[i] * w;
        }
    }
}

/*
 * oracle_bin_fine: Cells.add's binning into the fine (kappa, gamma) lattice
 * (sgl_full.py:1314-1383, 971-1046): gamma clipped below at 1.0000001 g_lo,
 * bins by searchsorted(edges, v, "right") with the last edge inside the last
 * bin, points outside the lattice dropped (their weight summed into *dropped);
 * weight, weight kappa and weight gamma accumulated in the points' order.
 */
static int64_t bin_right(const double *e, int64_t n, double v) {
    int64_t lo = 0, hi = n;               /* first index with e[i] > v */
    while (lo < hi) {
        int64_t mid = (lo + hi) >> 1;
        if (e[mid] > v) hi = mid; else lo = mid + 1;
    }
    return lo;
}

void oracle_bin_fine(int64_t n, const double *K, const double *G, const double *W,
                     const double *ek, int64_t nek, const double *eg, int64_t neg,
                     double *H0, double *H0k, double *H0g, double *dropped) {
    int64_t nkf = nek - 1, ngf = neg - 1;
    double drop = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double k = K[i], g = G[i], w = W[i];
        double gl = 1.0000001 * eg[0];
        if (g < gl) g = gl;
        int64_t a = bin_right(ek, nek, k), b = bin_right(eg, neg, g);
        if (k == ek[nek - 1]) a -= 1;
        if (g == eg[neg - 1]) b -= 1;
        if (a >= 1 && a <= nkf && b >= 1 && b <= ngf) {
            int64_t f = (a - 1) * ngf + (b - 1);
            H0[f] += w;
            H0k[f] += w * k;
            H0g[f] += w * g;
        } else {
            drop += w;
        }
    }
    *dropped += drop;
}

/* ------------------------------------------------------------------------
 * xi_lin deposit (sgl_full.py:1081-1130 _xi_add_grouped, 1142-1183
 * _bessel_pack_nb, 1397-1441 Cells._xi_deposit) for one row's points.
 * Points inside the coarse (kappa, gamma) box are grouped by their lens
 * distance R (stable order); per distinct R the table J_{2m}(k_j R), m = 0..2
 * (m >= 1 only for tangential shear), by the power series below x = 2 and by
 * J0, J1 and the upward recurrence above; then
 *   Hx[bin][m][j] += (sum of beta w of the group's points in the bin) J_{2m}(k_j R),
 *   lx[0][j] += (sum beta w kappa) J0(k_j R),  lx[1][j] += (sum beta w gamma) J2(k_j R).
 * Hx is [nkc * ngc][3][nkp], lx [2][nkp].
 */
static const double *xi_sort_R;
static int xi_cmp(const void *a, const void *b) {
    int64_t i = *(const int64_t *)a, j = *(const int64_t *)b;
    double ri = xi_sort_R[i], rj = xi_sort_R[j];
    if (ri < rj) return -1;
    if (ri > rj) return 1;
    return (i < j) ? -1 : (i > j);
}

static void xi_bessel(double x, int tang, double *J0, double *J2, double *J4) {
    if (x < 2.0) {
        double h = 0.5 * x, h2 = h * h;
        double t0 = 1.0, s0 = 1.0, t2 = 0.5, s2 = 0.5, t4 = 1.0 / 24.0, s4 = 1.0 / 24.0;
        int nt = (x < 0.02) ? 3 : ((x < 0.3) ? 6 : 14);
        for (int k = 1; k < nt; ++k) {
            double f = (-h2) / k;
            t0 *= f / k;
            s0 += t0;
            t2 *= f / (k + 2);
            s2 += t2;
            t4 *= f / (k + 4);
            s4 += t4;
        }
        *J0 = s0;
        *J2 = tang ? s2 * (h * h) : 0.0;
        *J4 = tang ? s4 * (h * h * h * h) : 0.0;
        return;
    }
    double j0v = j0(x);
    *J0 = j0v;
    if (!tang) { *J2 = 0.0; *J4 = 0.0; return; }
    double inv = 1.0 / x, j1v = j1(x);
    double j2v = 2.0 * inv * j1v - j0v;
    double j3v = 4.0 * inv * j2v - j1v;
    *J2 = j2v;
    *J4 = 6.0 * inv * j3v - j2v;
}

void oracle_xi_deposit(int64_t n, const double *K, const double *G, const double *W,
                       const double *R, double beta, int64_t tang, int64_t nkp, const double *kp,
                       const double *ekc, int64_t nek, const double *egc, int64_t neg,
                       double *Hx, double *lx) {
    int64_t nkc = nek - 1, ngc = neg - 1;
    int64_t *ord = malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    int64_t *bins = malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    double *slot = calloc((size_t)(nkc * ngc), sizeof(double));
    int64_t *mark = malloc((size_t)(nkc * ngc) * sizeof(int64_t));
    int64_t *touched = malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    double *Jt = malloc((size_t)(3 * nkp) * sizeof(double));
    for (int64_t f = 0; f < nkc * ngc; ++f) mark[f] = -1;
    for (int64_t i = 0; i < n; ++i) {
        ord[i] = i;
        double k = K[i], g = G[i];
        int64_t a = bin_right(ekc, nek, k), b = bin_right(egc, neg, g);
        if (k == ekc[nek - 1]) a -= 1;
        if (g == egc[neg - 1]) b -= 1;
        bins[i] = (a >= 1 && a <= nkc && b >= 1 && b <= ngc) ? (a - 1) * ngc + (b - 1) : -1;
    }
    xi_sort_R = R;
    qsort(ord, (size_t)n, sizeof(int64_t), xi_cmp);
    int64_t t = 0, u = 0;
    while (t < n) {
        int64_t t1 = t;
        double r = R[ord[t]];
        while (t1 < n && R[ord[t1]] == r) ++t1;
        int64_t nt = 0;
        double sk = 0.0, sg = 0.0;
        for (int64_t q = t; q < t1; ++q) {
            int64_t i = ord[q], f = bins[i];
            if (f < 0) continue;
            double bw = beta * W[i];
            if (mark[f] != u) { mark[f] = u; touched[nt++] = f; }
            slot[f] += bw;
            sk += bw * K[i];
            sg += bw * G[i];
        }
        if (nt > 0) {
            for (int64_t j = 0; j < nkp; ++j)
                xi_bessel(kp[j] * r, (int)tang, &Jt[j], &Jt[nkp + j], &Jt[2 * nkp + j]);
            for (int64_t q = 0; q < nt; ++q) {
                int64_t f = touched[q];
                double v = slot[f];
                slot[f] = 0.0;
                double *H = Hx + f * 3 * nkp;
                for (int m = 0; m < 3; ++m)
                    for (int64_t j = 0; j < nkp; ++j) H[m * nkp + j] += v * Jt[m * nkp + j];
            }
            for (int64_t j = 0; j < nkp; ++j) lx[j] += sk * Jt[j];
            for (int64_t j = 0; j < nkp; ++j) lx[nkp + j] += sg * Jt[nkp + j];
        }
        ++u;
        t = t1;
    }
    free(ord); free(bins); free(slot); free(mark); free(touched); free(Jt);
}

/* ------------------------------------------------------------------------
 * oracle_xi_rate: the positive closure's lens rate per coarse (kappa, gamma)
 * bin of one row's xi deposit points (_xi_rate_deposit, sgl_full.py:1270-1280):
 * the xi kernels' binning (searchsorted right, the last edge inside the last
 * bin), points outside the box dropped, Wb[bin] += w in the points' order.
 */
void oracle_xi_rate(int64_t n, const double *K, const double *G, const double *W,
                    const double *ekc, int64_t nek, const double *egc, int64_t neg, double *Wb) {
    int64_t nkc = nek - 1, ngc = neg - 1;
    for (int64_t i = 0; i < n; ++i) {
        double k = K[i], g = G[i];
        int64_t a = bin_right(ekc, nek, k), b = bin_right(egc, neg, g);
        if (k == ekc[nek - 1]) a -= 1;
        if (g == egc[neg - 1]) b -= 1;
        if (a >= 1 && a <= nkc && b >= 1 && b <= ngc) Wb[(a - 1) * ngc + (b - 1)] += W[i];
    }
}

/* ------------------------------------------------------------------------
 * oracle_vp_lnr: ln R of the positive closure on the tensor grid kv x qv
 * (FusedVP.lnR_blocked / _fused_groups, pos_closure_fused.py:52-140, and
 * _chain_recursion, pos_closure.py:183-240).
 *
 * Cells of a shell are grouped (gstart over cells, sh_g over groups); cell c
 * has the unique (kappa, gamma) index cu[c] and weight cw[c].  Per point p =
 * (k_a, q_b) and group g
 *   Y_g = sum_{c in g} w_c [(e^{i k kappa_u} - 1) J0(q gamma_u) + (J0 - 1)],
 * e - 1 as numpy's complex expm1 ((-2 sin^2(x/2), sin x)), J0 - 1 by plain
 * subtraction, as the reference forms them; the rates enter through their
 * interval form (rate_intervals: D_j = c + a (u_j - anchor) + b on the active
 * states lo..hi, c elsewhere): linear groups (c = 0) by the two sums
 * u_j sum a Y + sum b Y, clipped groups (c = -1) anchored at their first
 * (hi = ns - 1) or last active state, summed by prefix over the states.
 * Then per point the log-scaled forward recursion over the shells with the
 * transitions K[s] ([ns][ns], row i = from-state): alpha_0 = w e^{z_0},
 * alpha_s = (K[s-1]^T alpha_{s-1}) e^{z_s}, rescaled by max Re z and the
 * largest |alpha| each shell; ln R = scale + ln |sum alpha| + i arg(sum alpha).
 */
void oracle_vp_lnr(int64_t nsh, const int64_t *sh_g, const int64_t *gstart, const int64_t *cu,
                   const double *cw, const double *aa, const double *bb, const double *cc,
                   const int64_t *lo, const int64_t *hi, int64_t ns, const double *u,
                   const double *K, const double *w, int64_t nU, const double *ku,
                   const double *gu, int64_t nk, const double *kv, int64_t nq, const double *qv,
                   double *out_re, double *out_im) {
    const int64_t nkb = 4;
    double *J = malloc((size_t)(nU * nq) * sizeof(double));
    double *Jm1 = malloc((size_t)(nU * nq) * sizeof(double));
    double *er = malloc((size_t)(nU * nkb) * sizeof(double));
    double *ei = malloc((size_t)(nU * nkb) * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int64_t x = 0; x < nU; ++x)
        for (int64_t b = 0; b < nq; ++b) {
            double v = j0(gu[x] * qv[b]);
            J[x * nq + b] = v;
            Jm1[x * nq + b] = v - 1.0;
        }
    const int64_t Pmax = nkb * nq;
    double *zr = malloc((size_t)(nsh * ns * Pmax) * sizeof(double));
    double *zi = malloc((size_t)(nsh * ns * Pmax) * sizeof(double));
    for (int64_t a0 = 0; a0 < nk; a0 += nkb) {
        const int64_t na = (nk - a0 < nkb) ? nk - a0 : nkb, P = na * nq;
#pragma omp parallel for schedule(static)
        for (int64_t x = 0; x < nU; ++x)
            for (int64_t a = 0; a < na; ++a) {
                double y = kv[a0 + a] * ku[x], h = sin(0.5 * y);
                er[x * nkb + a] = -2.0 * h * h;
                ei[x * nkb + a] = sin(y);
            }
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t s = 0; s < nsh; ++s) {
            double *buf = calloc((size_t)(8 * ns * P + 6 * P + 2 * P + nq), sizeof(double));
            double *ar = buf, *ai = ar + ns * P, *br = ai + ns * P, *bi = br + ns * P;
            double *nar = bi + ns * P, *nai = nar + ns * P, *nbr = nai + ns * P, *nbi = nbr + ns * P;
            double *la = nbi + ns * P, *lai = la + P, *lb = lai + P, *lbi = lb + P;
            double *cr = lbi + P, *ci = cr + P, *accr = ci + P, *acci = accr + P, *S = acci + P;
            for (int64_t g = sh_g[s]; g < sh_g[s + 1]; ++g) {
                if (aa[g] == 0.0 && bb[g] == 0.0 && cc[g] == 0.0) continue;
                for (int64_t p = 0; p < P; ++p) { accr[p] = 0.0; acci[p] = 0.0; }
                for (int64_t b = 0; b < nq; ++b) S[b] = 0.0;
                for (int64_t c = gstart[g]; c < gstart[g + 1]; ++c) {
                    const int64_t x = cu[c];
                    const double wc = cw[c];
                    const double *Jx = J + x * nq, *Jm = Jm1 + x * nq;
                    for (int64_t b = 0; b < nq; ++b) S[b] += wc * Jm[b];
                    for (int64_t a = 0; a < na; ++a) {
                        const double xr = wc * er[x * nkb + a], xi = wc * ei[x * nkb + a];
                        double *Ar = accr + a * nq, *Ai = acci + a * nq;
                        for (int64_t b = 0; b < nq; ++b) {
                            Ar[b] += xr * Jx[b];
                            Ai[b] += xi * Jx[b];
                        }
                    }
                }
                for (int64_t a = 0; a < na; ++a)
                    for (int64_t b = 0; b < nq; ++b) {
                        const int64_t p = a * nq + b;
                        const double X = accr[p] + S[b], Yv = acci[p];
                        const double ax = aa[g] * X, ay = aa[g] * Yv, bx = bb[g] * X, by = bb[g] * Yv;
                        if (cc[g] == 0.0) {
                            la[p] += ax; lai[p] += ay; lb[p] += bx; lbi[p] += by;
                        } else if (hi[g] == ns - 1) {
                            const int64_t l = lo[g];
                            ar[l * P + p] += ax; ai[l * P + p] += ay;
                            br[l * P + p] += bx; bi[l * P + p] += by;
                            cr[p] -= X; ci[p] -= Yv;
                        } else {
                            const int64_t r = hi[g];
                            nar[r * P + p] += ax; nai[r * P + p] += ay;
                            nbr[r * P + p] += bx; nbi[r * P + p] += by;
                            cr[p] -= X; ci[p] -= Yv;
                        }
                    }
            }
            double *Zr = zr + s * ns * P, *Zi = zi + s * ns * P;
            for (int64_t p = 0; p < P; ++p) {
                double sx = 0.0, sy = 0.0, vx = 0.0, vy = 0.0;
                for (int64_t j = 0; j < ns; ++j) {
                    if (j > 0) { const double du = u[j] - u[j - 1]; vx += du * sx; vy += du * sy; }
                    vx += br[j * P + p]; vy += bi[j * P + p];
                    sx += ar[j * P + p]; sy += ai[j * P + p];
                    Zr[j * P + p] = u[j] * la[p] + lb[p] + vx + cr[p];
                    Zi[j * P + p] = u[j] * lai[p] + lbi[p] + vy + ci[p];
                }
                sx = sy = vx = vy = 0.0;
                for (int64_t j = ns - 1; j >= 0; --j) {
                    if (j < ns - 1) { const double du = u[j] - u[j + 1]; vx += du * sx; vy += du * sy; }
                    vx += nbr[j * P + p]; vy += nbi[j * P + p];
                    sx += nar[j * P + p]; sy += nai[j * P + p];
                    Zr[j * P + p] += vx;
                    Zi[j * P + p] += vy;
                }
            }
            free(buf);
        }
#pragma omp parallel for schedule(static)
        for (int64_t p = 0; p < P; ++p) {
            double arr[64], aii[64], brr[64], bii[64];
            double scale = 0.0;
            for (int64_t s = 0; s < nsh; ++s) {
                const double *Zr = zr + s * ns * P, *Zi = zi + s * ns * P;
                if (s == 0) {
                    for (int64_t j = 0; j < ns; ++j) { brr[j] = w[j]; bii[j] = 0.0; }
                } else {
                    const double *Ks = K + (s - 1) * ns * ns;
                    for (int64_t j = 0; j < ns; ++j) {
                        double tr = 0.0, ti = 0.0;
                        for (int64_t i = 0; i < ns; ++i) {
                            tr += Ks[i * ns + j] * arr[i];
                            ti += Ks[i * ns + j] * aii[i];
                        }
                        brr[j] = tr; bii[j] = ti;
                    }
                }
                double mx = Zr[p];
                for (int64_t j = 1; j < ns; ++j) if (Zr[j * P + p] > mx) mx = Zr[j * P + p];
                double nrm = 0.0;
                for (int64_t j = 0; j < ns; ++j) {
                    const double m = exp(Zr[j * P + p] - mx);
                    const double e_r = m * cos(Zi[j * P + p]), e_i = m * sin(Zi[j * P + p]);
                    arr[j] = brr[j] * e_r - bii[j] * e_i;
                    aii[j] = brr[j] * e_i + bii[j] * e_r;
                    const double a2 = arr[j] * arr[j] + aii[j] * aii[j];
                    if (a2 > nrm) nrm = a2;
                }
                nrm = sqrt(nrm);
                for (int64_t j = 0; j < ns; ++j) { arr[j] /= nrm; aii[j] /= nrm; }
                scale += mx + log(nrm);
            }
            double sr = 0.0, si = 0.0;
            for (int64_t j = 0; j < ns; ++j) { sr += arr[j]; si += aii[j]; }
            const int64_t a = p / nq, b = p - a * nq;
            out_re[(a0 + a) * nq + b] = scale + 0.5 * log(sr * sr + si * si);
            out_im[(a0 + a) * nq + b] = atan2(si, sr);
        }
    }
    free(J); free(Jm1); free(er); free(ei); free(zr); free(zi);
}
