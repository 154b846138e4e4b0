/*
 * CPU oracle for the lattice-Boltzmann step path -- TEST INFRASTRUCTURE ONLY.
 *
 * A float64 restatement, in plain C with OpenMP loops, of the reference's
 * per-step array functions.  The Python composer in oracle/oracle.py calls
 * these in the reference's order (aero/lbm/solver.py:276-464,
 * aero/lbm/solver3d.py:325-541).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker
 * or the timed CPU baseline; the product never links it.
 *
 * Layout everywhere is the reference's: populations f[q][z][y][x] with x
 * fastest (C-contiguous (Q, Nz, Ny, Nx); 2D uses nz = 1 and Q = 9).
 * Lattice tables are passed in by the caller (ex, ey, ez, w, opp), so one
 * code path serves D2Q9 and D3Q19 exactly as the reference's generic loops do.
 *
 * Pinning: every function here is checked against the golden vectors written
 * from the reference's own NumPy path (tests/golden/make_golden.py) and against
 * closed-form invariants in tests/test_oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IDX(q, z, y, x) ((((size_t)(q) * nz + (z)) * ny + (y)) * nx + (x))
#define CELL(z, y, x) ((((size_t)(z)) * ny + (y)) * nx + (x))

static inline double feq_i(double w, double rho, double eu, double usq) {
    /* aero/lbm/d2q9.py:87, aero/lbm/d3q19.py:217 */
    return w * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq);
}

/* ------------------------------------------------------------------------ */
/* Collision: BGK (kernels.py:42-92), TRT (kernels_trt.py:183-217),         */
/* MRT (kernels_mrt.py:93-134, kernels3d_mrt.py:228-263).                   */
/* coll: 0 = BGK, 1 = TRT, 2 = MRT.  omega_field may be NULL.               */
/* ------------------------------------------------------------------------ */
int orc_collide(int Q, int nz, int ny, int nx,
                const double *f, double *fpost, const uint8_t *solid,
                const int *ex, const int *ey, const int *ez, const double *w,
                const int *opp, int coll, double omega, const double *omega_field,
                double trt_lambda, const double *M, const double *Minv,
                const double *s, int trt_variant) {
    const size_t N = (size_t)nz * ny * nx;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)N; ++c) {
        double fl[19], fe[19];
        double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
        for (int i = 0; i < Q; ++i) {
            double fi = f[(size_t)i * N + c];
            fl[i] = fi;
            rho += fi;
            mx += ex[i] * fi;
            my += ey[i] * fi;
            mz += ez[i] * fi;
        }
        double inv = rho > 0.0 ? 1.0 / rho : 0.0;
        double ux = inv * mx, uy = inv * my, uz = inv * mz;
        if (solid[c]) { ux = 0.0; uy = 0.0; uz = 0.0; }
        double usq = ux * ux + uy * uy + uz * uz;
        for (int i = 0; i < Q; ++i) {
            double eu = ex[i] * ux + ey[i] * uy + ez[i] * uz;
            fe[i] = feq_i(w[i], rho, eu, usq);
        }
        double om = omega_field ? omega_field[c] : omega;
        if (coll == 0) {
            for (int i = 0; i < Q; ++i)
                fpost[(size_t)i * N + c] = (1.0 - om) * fl[i] + om * fe[i];
        } else if (coll == 1) {
            /* trt2d.py:8-18 and kernels_trt.py:14-18: s- = 1/tau-,
               tau+ = lam*tau + (1-lam), tau- = (tau+ tau - 1/2)/(tau+ - 1/2) */
            double tau = 1.0 / om;
            double sm;
            if (trt_variant == 1) {
                /* standard magic-parameter TRT (not the reference operator):
                   tau+ = tau, tau- = Lambda/(tau - 1/2) + 1/2 */
                sm = 1.0 / (trt_lambda / (tau - 0.5) + 0.5);
            } else {
                double tp = trt_lambda * tau + (1.0 - trt_lambda);
                double tm = (tp * tau - 0.5) / (tp - 0.5);
                sm = 1.0 / tm;
            }
            for (int i = 0; i < Q; ++i) {
                int j = opp[i];
                double fp = 0.5 * (fl[i] + fl[j]);
                double fm = 0.5 * (fl[i] - fl[j]);
                double ep = 0.5 * (fe[i] + fe[j]);
                double em = 0.5 * (fe[i] - fe[j]);
                fpost[(size_t)i * N + c] = fe[i] + (1.0 - om) * (fp - ep) + (1.0 - sm) * (fm - em);
            }
        } else {
            double m[19], meq[19], mp[19];
            for (int a = 0; a < Q; ++a) {
                double acc = 0.0, acce = 0.0;
                for (int b = 0; b < Q; ++b) {
                    acc += M[a * Q + b] * fl[b];
                    acce += M[a * Q + b] * fe[b];
                }
                m[a] = acc;
                meq[a] = acce;
            }
            for (int a = 0; a < Q; ++a) mp[a] = m[a] - s[a] * (m[a] - meq[a]);
            for (int i = 0; i < Q; ++i) {
                double acc = 0.0;
                for (int a = 0; a < Q; ++a) acc += Minv[i * Q + a] * mp[a];
                fpost[(size_t)i * N + c] = acc;
            }
        }
    }
    return 0;
}

/* Periodic push stream: dst[i, (z+ez)%nz, (y+ey)%ny, (x+ex)%nx] = src[i,z,y,x]
   (kernels.py:94-115, kernels3d.py:63-80). */
int orc_stream(int Q, int nz, int ny, int nx, const double *src, double *dst,
               const int *ex, const int *ey, const int *ez) {
    for (int i = 0; i < Q; ++i) {
#pragma omp parallel for schedule(static)
        for (int z = 0; z < nz; ++z) {
            int zn = ((z + ez[i]) % nz + nz) % nz;
            for (int y = 0; y < ny; ++y) {
                int yn = ((y + ey[i]) % ny + ny) % ny;
                for (int x = 0; x < nx; ++x) {
                    int xn = ((x + ex[i]) % nx + nx) % nx;
                    dst[IDX(i, zn, yn, xn)] = src[IDX(i, z, y, x)];
                }
            }
        }
    }
    return 0;
}

/* Link list (build_surface_links*, boundary.py:232-280, boundary3d.py:279-330):
   cell fluid and its CLIPPED neighbour along e_i solid; ordered by i, then
   row-major cell.  Returns the count; writes up to cap entries of
   (i, z, y, x) into links (4 ints each).  Pass links = NULL to count. */
long long orc_build_links(int Q, int nz, int ny, int nx, const uint8_t *solid,
                          const int *ex, const int *ey, const int *ez,
                          int *links, long long cap) {
    long long n = 0;
    for (int i = 1; i < Q; ++i) {
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y)
                for (int x = 0; x < nx; ++x) {
                    if (solid[CELL(z, y, x)]) continue;
                    int zn = z + ez[i], yn = y + ey[i], xn = x + ex[i];
                    zn = zn < 0 ? 0 : (zn > nz - 1 ? nz - 1 : zn);
                    yn = yn < 0 ? 0 : (yn > ny - 1 ? ny - 1 : yn);
                    xn = xn < 0 ? 0 : (xn > nx - 1 ? nx - 1 : xn);
                    if (!solid[CELL(zn, yn, xn)]) continue;
                    if (links && n < cap) {
                        links[4 * n + 0] = i;
                        links[4 * n + 1] = z;
                        links[4 * n + 2] = y;
                        links[4 * n + 3] = x;
                    }
                    ++n;
                }
    }
    return n;
}

/* Bounce-back (boundary.py:283-300) and Bouzidi (boundary.py:303-352,
   boundary3d.py:353-393).  links: (L, 4) = (i, z, y, x). q may be NULL for
   plain mid-link bounce-back.  Reads only fpre, so link order is immaterial. */
int orc_bounce_back(int Q, int nz, int ny, int nx, double *f, const double *fpre,
                    long long L, const int *links, const float *q,
                    const int *ex, const int *ey, const int *ez, const int *opp, int variant) {
    (void)Q;
#pragma omp parallel for schedule(static)
    for (long long k = 0; k < L; ++k) {
        int i = links[4 * k], z = links[4 * k + 1], y = links[4 * k + 2], x = links[4 * k + 3];
        int o = opp[i];
        if (!q) {
            f[IDX(o, z, y, x)] = fpre[IDX(i, z, y, x)];
            continue;
        }
        double qq = (double)q[k];
        int zn = z + ez[o], yn = y + ey[o], xn = x + ex[o];
        zn = zn < 0 ? 0 : (zn > nz - 1 ? nz - 1 : zn);
        yn = yn < 0 ? 0 : (yn > ny - 1 ? ny - 1 : yn);
        xn = xn < 0 ? 0 : (xn > nx - 1 ? nx - 1 : xn);
        /* variant 1: the reference rule; variant 2: linear Bouzidi-Firdaouss-
           Lallemand (2001), which takes the upstream partner f_i(x - e_i) for
           q < 1/2 and f_opp(x) for q >= 1/2 (not a reference operator) */
        int upstream = variant == 2 ? (qq < 0.5) : !(qq < 0.5);
        double partner = upstream ? fpre[IDX(i, zn, yn, xn)] : fpre[IDX(o, z, y, x)];
        if (qq < 0.5) {
            f[IDX(o, z, y, x)] = 2.0 * qq * fpre[IDX(i, z, y, x)] + (1.0 - 2.0 * qq) * partner;
        } else {
            double inv2q = 1.0 / (2.0 * qq);
            f[IDX(o, z, y, x)] = inv2q * fpre[IDX(i, z, y, x)] + (1.0 - inv2q) * partner;
        }
    }
    return 0;
}

/* ---------------------------- 2D inlets --------------------------------- */
/* Zou-He velocity inlet with optional travelling uy perturbation
   (boundary.py:15-45). */
int orc_inlet_zouhe_2d(int ny, int nx, double *f, double u0, double uy_amp, long long step) {
    const int nz = 1;
    for (int y = 0; y < ny; ++y) {
#define F(i) f[IDX(i, 0, y, 0)]
        double rho = ((F(0) + F(2) + F(4)) + 2.0 * (F(3) + F(6) + F(7))) / (1.0 - u0);
        F(1) = F(3) + (2.0 / 3.0) * rho * u0;
        F(5) = F(7) - 0.5 * (F(2) - F(4)) + (1.0 / 6.0) * rho * u0;
        F(8) = F(6) + 0.5 * (F(2) - F(4)) + (1.0 / 6.0) * rho * u0;
        if (uy_amp > 0.0) {
            double uy = uy_amp * u0 * sin(2.0 * M_PI * (double)y / (double)(ny > 1 ? ny : 1) + 0.17 * (double)step);
            F(2) += uy * 0.25;
            F(4) -= uy * 0.25;
            F(5) += uy * 0.15;
            F(8) -= uy * 0.15;
        }
#undef F
    }
    return 0;
}

/* Per-row target velocities (boundary.py:48-65). */
int orc_inlet_velocity_field_2d(int ny, int nx, double *f, const double *ux_t, const double *uy_t) {
    const int nz = 1;
    for (int y = 0; y < ny; ++y) {
#define F(i) f[IDX(i, 0, y, 0)]
        double den = 1.0 - ux_t[y];
        if (den < 1e-8) den = 1e-8;
        double rho = ((F(0) + F(2) + F(4)) + 2.0 * (F(3) + F(6) + F(7))) / den;
        double cy = rho * uy_t[y];
        F(1) = F(3) + (2.0 / 3.0) * rho * ux_t[y];
        F(5) = F(7) - 0.5 * cy + (1.0 / 6.0) * rho * ux_t[y];
        F(8) = F(6) + 0.5 * cy + (1.0 / 6.0) * rho * ux_t[y];
#undef F
    }
    return 0;
}

/* Zou-He pressure inlet / outlet (boundary.py:116-159). */
int orc_inlet_pressure_2d(int ny, int nx, double *f, double rho_in) {
    const int nz = 1;
    for (int y = 0; y < ny; ++y) {
#define F(i) f[IDX(i, 0, y, 0)]
        double ux = 1.0 - (F(0) + F(2) + F(4) + 2.0 * (F(3) + F(6) + F(7))) / rho_in;
        F(1) = F(3) + (2.0 / 3.0) * rho_in * ux;
        F(5) = F(7) - 0.5 * (F(2) - F(4)) + (1.0 / 6.0) * rho_in * ux;
        F(8) = F(6) + 0.5 * (F(2) - F(4)) + (1.0 / 6.0) * rho_in * ux;
#undef F
    }
    return 0;
}

int orc_outlet_pressure_2d(int ny, int nx, double *f, double rho_out) {
    const int nz = 1;
    for (int y = 0; y < ny; ++y) {
#define F(i) f[IDX(i, 0, y, nx - 1)]
        double ux = -1.0 + (F(0) + F(2) + F(4) + 2.0 * (F(1) + F(5) + F(8))) / rho_out;
        F(3) = F(1) - (2.0 / 3.0) * rho_out * ux;
        F(7) = F(5) + 0.5 * (F(2) - F(4)) - (1.0 / 6.0) * rho_out * ux;
        F(6) = F(8) - 0.5 * (F(2) - F(4)) - (1.0 / 6.0) * rho_out * ux;
#undef F
    }
    return 0;
}

/* ---------------------------- 3D inlets --------------------------------- */
/* mode 0: Zou-He with uz/uy sine perturbation (boundary3d.py:24-68);
   mode 1: per-cell velocity field (boundary3d.py:71-93);
   mode 2: SEM fluctuations (boundary3d.py:96-137).
   Plane arrays are (nz, ny). */
int orc_inlet_3d(int nz, int ny, int nx, double *f, int mode, double u0,
                 double uy_amp, double uz_amp, long long step,
                 const double *pu, const double *pv, const double *pw) {
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
#define F(i) f[IDX(i, z, y, 0)]
            double Fm = F(2) + F(8) + F(10) + F(12) + F(14);
            double Fn = F(0) + F(3) + F(4) + F(5) + F(6) + F(15) + F(16) + F(17) + F(18);
            double cyb = F(3) - F(4) + F(15) - F(16) + F(17) - F(18);
            double czb = F(5) - F(6) + F(15) + F(16) - F(17) - F(18);
            double ux, rho, cy, cz;
            size_t p = (size_t)z * ny + y;
            if (mode == 0) {
                ux = u0;
                rho = (2.0 * Fm + Fn) / (1.0 - u0);
                cy = cyb;
                cz = czb;
                double phase = 0.17 * (double)step;
                if (uy_amp > 0.0)
                    cy = cy + rho * (uy_amp * u0 * sin(2.0 * M_PI * (double)y / (double)(ny > 1 ? ny : 1) + phase));
                if (uz_amp > 0.0)
                    cz = cz + rho * (uz_amp * u0 * sin(2.0 * M_PI * (double)z / (double)(nz > 1 ? nz : 1) + phase));
            } else if (mode == 1) {
                ux = pu[p];
                double den = 1.0 - ux;
                if (den < 1e-8) den = 1e-8;
                rho = (2.0 * Fm + Fn) / den;
                cy = rho * pv[p];
                cz = rho * pw[p];
            } else {
                ux = u0 + pu[p];
                double den = 1.0 - ux;
                if (den < 1e-10) den = 1e-10;
                rho = (2.0 * Fm + Fn) / den;
                cy = cyb - rho * pv[p];
                cz = czb - rho * pw[p];
            }
            F(1) = F(2) + (1.0 / 3.0) * rho * ux;
            F(7) = F(10) + (1.0 / 6.0) * rho * ux - 0.5 * cy;
            F(9) = F(8) + (1.0 / 6.0) * rho * ux + 0.5 * cy;
            F(11) = F(14) + (1.0 / 6.0) * rho * ux - 0.5 * cz;
            F(13) = F(12) + (1.0 / 6.0) * rho * ux + 0.5 * cz;
#undef F
        }
    return 0;
}

/* Recycling-rescaling inlet (boundary3d.py:172-198): plane x = nx-2 -> x = 0. */
int orc_recycling_3d(int nz, int ny, int nx, double *f, double target_u0,
                     const int *ex, const int *ey, const int *ez, const double *w) {
    const int Q = 19;
    const size_t P = (size_t)nz * ny;
    double *r = (double *)malloc(sizeof(double) * 4 * P);
    double sum_ux = 0.0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y) {
            double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
            for (int i = 0; i < Q; ++i) {
                double fi = f[IDX(i, z, y, nx - 2)];
                rho += fi;
                mx += ex[i] * fi;
                my += ey[i] * fi;
                mz += ez[i] * fi;
            }
            double inv = rho > 0.0 ? 1.0 / rho : 0.0;
            size_t p = (size_t)z * ny + y;
            r[4 * p] = rho;
            r[4 * p + 1] = inv * mx;
            r[4 * p + 2] = inv * my;
            r[4 * p + 3] = inv * mz;
            sum_ux += inv * mx;
        }
    double mean = sum_ux / (double)P;
    double scale = fabs(mean) < 1e-12 ? 1.0 : target_u0 / mean;
    for (size_t p = 0; p < P; ++p) {
        int z = (int)(p / ny), y = (int)(p % ny);
        double rho = r[4 * p], ux = r[4 * p + 1] * scale, uy = r[4 * p + 2] * scale, uz = r[4 * p + 3] * scale;
        double usq = ux * ux + uy * uy + uz * uz;
        for (int i = 0; i < Q; ++i) {
            double eu = ex[i] * ux + ey[i] * uy + ez[i] * uz;
            f[IDX(i, z, y, 0)] = feq_i(w[i], rho, eu, usq);
        }
    }
    free(r);
    return 0;
}

/* ---------------------------- outlets ----------------------------------- */
/* Zero-gradient (boundary.py:68-75) and convective (boundary.py:78-113);
   prev is (Q, nz, ny) and is updated in place for the convective case. */
int orc_outlet_zero_gradient(int Q, int nz, int ny, int nx, double *f) {
    for (int i = 0; i < Q; ++i)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) f[IDX(i, z, y, nx - 1)] = f[IDX(i, z, y, nx - 2)];
    return 0;
}

int orc_outlet_convective(int Q, int nz, int ny, int nx, double *f, double *prev, double u_conv) {
    double u = u_conv < 0.0 ? 0.0 : (u_conv > 1.0 ? 1.0 : u_conv);
    for (int i = 0; i < Q; ++i)
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) {
                size_t p = ((size_t)i * nz + z) * ny + y;
                double v = (1.0 - u) * prev[p] + u * f[IDX(i, z, y, nx - 2)];
                f[IDX(i, z, y, nx - 1)] = v;
                prev[p] = v;
            }
    return 0;
}

/* Sponge (sponge.py:8-70): relax toward feq(rho_local, u_target, 0). */
int orc_sponge(int Q, int nz, int ny, int nx, double *f, const double *sigma, double u_t,
               const int *ex, const double *w) {
    for (int x = 0; x < nx; ++x) {
        double s = sigma[x];
        if (s <= 0.0) continue;
        for (int z = 0; z < nz; ++z)
            for (int y = 0; y < ny; ++y) {
                double rho = 0.0;
                for (int i = 0; i < Q; ++i) rho += f[IDX(i, z, y, x)];
                double usq = u_t * u_t;
                for (int i = 0; i < Q; ++i) {
                    double eu = ex[i] * u_t;
                    double fe = feq_i(w[i], rho, eu, usq);
                    f[IDX(i, z, y, x)] = (1.0 - s) * f[IDX(i, z, y, x)] + s * fe;
                }
            }
    }
    return 0;
}

/* Walls at y = 0 / y = ny-1 (boundary.py:162-226, boundary3d.py:205-272).
   mode: 0 slip, 1 noslip, 2 moving.  dst/src pairs are given per wall by the
   caller (n pairs), and moving adds +/- 6 w rho u to the pairs flagged in sgn. */
int orc_walls(int Q, int nz, int ny, int nx, double *f, int npairs,
              const int *bot_dst, const int *bot_src, const int *bot_sgn,
              const int *top_dst, const int *top_src, const int *top_sgn,
              double u_top, double u_bottom) {
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x) {
            double rb = 0.0, rt = 0.0;
            for (int i = 0; i < Q; ++i) {
                rb += f[IDX(i, z, 0, x)];
                rt += f[IDX(i, z, ny - 1, x)];
            }
            rb = rb > 1e-12 ? rb : 1e-12;
            rt = rt > 1e-12 ? rt : 1e-12;
            for (int k = 0; k < npairs; ++k)
                f[IDX(bot_dst[k], z, 0, x)] = f[IDX(bot_src[k], z, 0, x)] + bot_sgn[k] * (6.0 * (1.0 / 36.0) * rb * u_bottom);
            for (int k = 0; k < npairs; ++k)
                f[IDX(top_dst[k], z, ny - 1, x)] = f[IDX(top_src[k], z, ny - 1, x)] + top_sgn[k] * (6.0 * (1.0 / 36.0) * rt * u_top);
        }
    return 0;
}

/* Uniform Guo-style body force on fluid cells (ibm_guo.py:230-301). */
int orc_body_force(int Q, int nz, int ny, int nx, double *f, const uint8_t *solid, double tau,
                   double fx, double fy, double fz, const int *ex, const int *ey, const int *ez,
                   const double *w) {
    const size_t N = (size_t)nz * ny * nx;
    const double cs2 = 1.0 / 3.0;
    const double coeff = 1.0 - 0.5 / tau;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)N; ++c) {
        if (solid[c]) continue;
        double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
        for (int i = 0; i < Q; ++i) {
            double fi = f[(size_t)i * N + c];
            rho += fi;
            mx += ex[i] * fi;
            my += ey[i] * fi;
            mz += ez[i] * fi;
        }
        if (rho <= 0.0) continue;
        double ux = mx / rho, uy = my / rho, uz = mz / rho;
        double uf = ux * fx + uy * fy + uz * fz;
        for (int i = 0; i < Q; ++i) {
            double eu = ex[i] * ux + ey[i] * uy + ez[i] * uz;
            double ef = ex[i] * fx + ey[i] * fy + ez[i] * fz;
            double term = (ef - uf) / cs2 + (eu * ef) / (cs2 * cs2);
            f[(size_t)i * N + c] += coeff * w[i] * rho * term;
        }
    }
    return 0;
}

/* Guo immersed-boundary forcing toward a resting wall (ibm_guo.py:110-182):
   on fluid cells with 0 < phi <= 1.5 (a NaN phi is not skipped, as there),
   F = (1 - 1/(2 tau)) (0 - u) and f_i += w_i e_i.F, with rho and u of the
   cell's own populations (skipped where rho <= 0). */
int orc_ibm_guo(int Q, int nz, int ny, int nx, double *f, const double *phi, const uint8_t *solid,
                double tau, const int *ex, const int *ey, const int *ez, const double *w) {
    const size_t N = (size_t)nz * ny * nx;
    const double delta = 1.5;
    const double coeff = 1.0 - 0.5 / tau;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)N; ++c) {
        if (solid[c] || phi[c] <= 0.0 || phi[c] > delta) continue;
        double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
        for (int i = 0; i < Q; ++i) {
            double fi = f[(size_t)i * N + c];
            rho += fi;
            mx += ex[i] * fi;
            my += ey[i] * fi;
            mz += ez[i] * fi;
        }
        if (rho <= 0.0) continue;
        double Fx = coeff * (0.0 - mx / rho), Fy = coeff * (0.0 - my / rho), Fz = coeff * (0.0 - mz / rho);
        for (int i = 0; i < Q; ++i)
            f[(size_t)i * N + c] += w[i] * (ex[i] * Fx + ey[i] * Fy + ez[i] * Fz);
    }
    return 0;
}

/* Macroscopic fields with the rho > 0 guard (d2q9.py:44-63); u is not zeroed
   in solids. Outputs (N,) arrays; uz may be NULL. */
int orc_macroscopic(int Q, int nz, int ny, int nx, const double *f, const int *ex, const int *ey,
                    const int *ez, double *rho_o, double *ux_o, double *uy_o, double *uz_o) {
    const size_t N = (size_t)nz * ny * nx;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)N; ++c) {
        double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
        for (int i = 0; i < Q; ++i) {
            double fi = f[(size_t)i * N + c];
            rho += fi;
            mx += ex[i] * fi;
            my += ey[i] * fi;
            mz += ez[i] * fi;
        }
        double inv = rho > 0.0 ? 1.0 / rho : 0.0;
        rho_o[c] = rho;
        ux_o[c] = inv * mx;
        uy_o[c] = inv * my;
        if (uz_o) uz_o[c] = inv * mz;
    }
    return 0;
}

/* LES relaxation field (les.py:273-344) from velocity fields: central
   differences on interior fluid cells; model 0 Smagorinsky, 1 WALE;
   damp (N,) or NULL is the Van Driest factor multiplying C_s. */
int orc_les_omega(int dim, int nz, int ny, int nx, const double *ux, const double *uy,
                  const double *uz, const uint8_t *solid, double base_nu, double cs, int model,
                  const double *damp, double *omega_out) {
    const size_t N = (size_t)nz * ny * nx;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)N; ++c) {
        int x = (int)(c % nx), y = (int)((c / nx) % ny), z = (int)(c / ((size_t)nx * ny));
        double nu_sgs = 0.0;
        int interior = (x >= 1 && x <= nx - 2 && y >= 1 && y <= ny - 2);
        if (dim == 3) interior = interior && z >= 1 && z <= nz - 2;
        if (interior && !solid[c]) {
            double csl = damp ? cs * damp[c] : cs;
            double g[3][3] = {{0}};
            const double *u[3] = {ux, uy, uz};
            int nd = dim;
            for (int a = 0; a < nd; ++a) {
                g[a][0] = 0.5 * (u[a][c + 1] - u[a][c - 1]);
                g[a][1] = 0.5 * (u[a][c + nx] - u[a][c - nx]);
                if (dim == 3) g[a][2] = 0.5 * (u[a][c + (size_t)nx * ny] - u[a][c - (size_t)nx * ny]);
            }
            if (model == 0) {
                double smag;
                if (dim == 2) {
                    double s11 = g[0][0], s22 = g[1][1], s12 = 0.5 * (g[0][1] + g[1][0]);
                    smag = sqrt(2.0 * (s11 * s11 + s22 * s22 + 2.0 * s12 * s12));
                } else {
                    double s11 = g[0][0], s22 = g[1][1], s33 = g[2][2];
                    double s12 = 0.5 * (g[0][1] + g[1][0]);
                    double s13 = 0.5 * (g[0][2] + g[2][0]);
                    double s23 = 0.5 * (g[1][2] + g[2][1]);
                    smag = sqrt(2.0 * (s11 * s11 + s22 * s22 + s33 * s33) + 4.0 * (s12 * s12 + s13 * s13 + s23 * s23));
                }
                nu_sgs = (csl * csl) * smag;
            } else {
                double s_sq, sd_sq;
                if (dim == 2) {
                    double g11 = g[0][0], g12 = g[0][1], g21 = g[1][0], g22 = g[1][1];
                    double s11 = g11, s22 = g22, s12 = 0.5 * (g12 + g21);
                    s_sq = s11 * s11 + s22 * s22 + 2.0 * s12 * s12;
                    double q11 = g11 * g11 + g12 * g21, q12 = g11 * g12 + g12 * g22;
                    double q21 = g21 * g11 + g22 * g21, q22 = g21 * g12 + g22 * g22;
                    double tr = 0.5 * (q11 + q22);
                    double d11 = q11 - tr, d22 = q22 - tr, d12 = 0.5 * (q12 + q21);
                    sd_sq = d11 * d11 + d22 * d22 + 2.0 * d12 * d12;
                } else {
                    double sq = 0.0;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            double sab = 0.5 * (g[a][b] + g[b][a]);
                            sq += sab * sab;
                        }
                    s_sq = sq;
                    double g2[3][3];
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            double acc = 0.0;
                            for (int k = 0; k < 3; ++k) acc += g[a][k] * g[k][b];
                            g2[a][b] = acc;
                        }
                    double tr = (g2[0][0] + g2[1][1] + g2[2][2]) / 3.0;
                    double sd = 0.0;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            double v = 0.5 * (g2[a][b] + g2[b][a]) - (a == b ? tr : 0.0);
                            sd += v * v;
                        }
                    sd_sq = sd;
                }
                double num = pow(sd_sq, 1.5);
                double den = pow(s_sq, 2.5) + pow(sd_sq, 1.25) + 1e-30;
                nu_sgs = (csl * csl) * num / den;
            }
        }
        double nu_eff = solid[c] ? base_nu : base_nu + nu_sgs;
        omega_out[c] = 1.0 / (3.0 * nu_eff + 0.5);
    }
    return 0;
}

/* Local Smagorinsky relaxation field from the populations f (Q x N) about to
   be collided (no reference counterpart: the relaxation-time closure of Hou,
   Sterling, Chen & Doolen 1996).  Per fluid cell, with Pi_ab = sum_i e_ia e_ib
   f_i, Pi^neq = Pi - rho (delta_ab / 3 + u_a u_b),
   Qn = sqrt(2 sum_ab Pi^neq_ab^2) and C = cs (times damp[c] when given):
       tau = (tau0 + sqrt(tau0^2 + 18 C^2 Qn / rho)) / 2,  omega_out = 1 / tau.
   Solid cells get omega.  The closure solves tau = tau0 + 3 C^2 |S| with the
   Chapman-Enskog strain |S| = 3 Qn / (2 rho tau) of the same populations. */
int orc_les_local_omega(int Q, int nz, int ny, int nx, const double *f, const uint8_t *solid,
                        const int *ex, const int *ey, const int *ez, double omega, double cs,
                        const double *damp, double *omega_out) {
    const size_t N = (size_t)nz * ny * nx;
    const double tau0 = 1.0 / omega;
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)N; ++c) {
        if (solid[c]) { omega_out[c] = omega; continue; }
        double rho = 0.0, m[3] = {0.0, 0.0, 0.0}, P[3][3] = {{0.0}};
        for (int i = 0; i < Q; ++i) {
            const double fi = f[(size_t)i * N + c];
            const int e[3] = {ex[i], ey[i], ez[i]};
            rho += fi;
            for (int a = 0; a < 3; ++a) {
                m[a] += e[a] * fi;
                for (int b = 0; b < 3; ++b) P[a][b] += e[a] * e[b] * fi;
            }
        }
        const double inv = rho > 0.0 ? 1.0 / rho : 0.0;
        const double u[3] = {m[0] * inv, m[1] * inv, m[2] * inv};
        const int dim = Q == 9 ? 2 : 3;
        double pp = 0.0;
        for (int a = 0; a < dim; ++a)
            for (int b = 0; b < dim; ++b) {
                const double v = P[a][b] - rho * ((a == b ? 1.0 / 3.0 : 0.0) + u[a] * u[b]);
                pp += v * v;
            }
        const double C = damp ? cs * damp[c] : cs;
        const double qn = sqrt(2.0 * pp);
        const double tau = 0.5 * (tau0 + sqrt(tau0 * tau0 + 18.0 * C * C * qn * inv));
        omega_out[c] = 1.0 / tau;
    }
    return 0;
}

/* Passive scalar + Boussinesq step (solver.py:427-443, solver3d.py:480-503,
   thermal.py): f gets the buoyancy force, g is collided, streamed and given
   the y-wall Dirichlet values.  gout must not alias g. */
int orc_thermal(int Q, int nz, int ny, int nx, double *f, const double *g, double *gout,
                const uint8_t *solid, const int *ex, const int *ey, const int *ez, const double *w,
                double omega_T, int buoyancy, double g_grav, double beta, double T_ref,
                double T_hot, double T_cold) {
    const size_t N = (size_t)nz * ny * nx;
    double *gp = (double *)malloc(sizeof(double) * Q * N);
#pragma omp parallel for schedule(static)
    for (long long c = 0; c < (long long)N; ++c) {
        double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0, T = 0.0;
        for (int i = 0; i < Q; ++i) {
            double fi = f[(size_t)i * N + c];
            rho += fi;
            mx += ex[i] * fi;
            my += ey[i] * fi;
            mz += ez[i] * fi;
            T += g[(size_t)i * N + c];
        }
        double inv = rho > 0.0 ? 1.0 / rho : 0.0;
        double ux = inv * mx, uy = inv * my, uz = inv * mz;
        if (buoyancy) {
            double buoy = g_grav * beta * (T - T_ref);
            for (int i = 0; i < Q; ++i) f[(size_t)i * N + c] += w[i] * 3.0 * rho * buoy * (double)ey[i];
        }
        for (int i = 0; i < Q; ++i) {
            double eu = ex[i] * ux + ey[i] * uy + ez[i] * uz;
            double geq = w[i] * T * (1.0 + 3.0 * eu);
            double gi = g[(size_t)i * N + c];
            gp[(size_t)i * N + c] = solid[c] ? w[i] * T : gi - omega_T * (gi - geq);
        }
    }
    orc_stream(Q, nz, ny, nx, gp, gout, ex, ey, ez);
    for (int i = 0; i < Q; ++i)
        for (int z = 0; z < nz; ++z)
            for (int x = 0; x < nx; ++x) {
                gout[IDX(i, z, 0, x)] = w[i] * T_hot;
                gout[IDX(i, z, ny - 1, x)] = w[i] * T_cold;
            }
    free(gp);
    return 0;
}

/* Momentum exchange (forces.py, forces3d.py).  out[12]:
   Fx Fy Fz  Mx My Mz  Fpx Fpy Fpz  Fvx Fvy Fvz (raw lattice units; 2D has
   Fz = Mx = My = 0 and Mz is the z-moment).  profile (nprof*3) gets fx, fy, fz
   summed per y (2D) or per z (3D); NULL skips it. */
int orc_forces(int Q, int nz, int ny, int nx, const double *fpre, const double *fpost,
               long long L, const int *links, const int *ex, const int *ey, const int *ez,
               const double *w, const int *opp, double cx, double cy, double cz, double *out,
               double *profile) {
    const size_t N = (size_t)nz * ny * nx;
    double acc[12] = {0};
    for (long long k = 0; k < L; ++k) {
        int i = links[4 * k], z = links[4 * k + 1], y = links[4 * k + 2], x = links[4 * k + 3];
        int o = opp[i];
        size_t c = CELL(z, y, x);
        double fo = fpre[(size_t)i * N + c], fi = fpost[(size_t)o * N + c];
        double mom = fo + fi;
        double dfx = ex[i] * mom, dfy = ey[i] * mom, dfz = ez[i] * mom;
        double rx = (double)x - cx, ry = (double)y - cy, rz = (double)z - cz;
        acc[0] += dfx;
        acc[1] += dfy;
        acc[2] += dfz;
        acc[3] += ry * dfz - rz * dfy;
        acc[4] += rz * dfx - rx * dfz;
        acc[5] += rx * dfy - ry * dfx;
        /* pressure/viscous split: feq from compute_macroscopic(f_pre[:, cell]) */
        double rho = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
        for (int j = 0; j < Q; ++j) {
            double fj = fpre[(size_t)j * N + c];
            rho += fj;
            mx += ex[j] * fj;
            my += ey[j] * fj;
            mz += ez[j] * fj;
        }
        double inv = rho > 0.0 ? 1.0 / rho : 0.0;
        double ux = inv * mx, uy = inv * my, uz = inv * mz;
        double usq = ux * ux + uy * uy + uz * uz;
        double feo = feq_i(w[i], rho, ex[i] * ux + ey[i] * uy + ez[i] * uz, usq);
        double fei = feq_i(w[o], rho, ex[o] * ux + ey[o] * uy + ez[o] * uz, usq);
        double mp = feo + fei;
        double mv = (fo - feo) + (fi - fei);
        acc[6] += ex[i] * mp;
        acc[7] += ey[i] * mp;
        acc[8] += ez[i] * mp;
        acc[9] += ex[i] * mv;
        acc[10] += ey[i] * mv;
        acc[11] += ez[i] * mv;
        if (profile) {
            int r = (nz == 1) ? y : z;
            profile[3 * r + 0] += dfx;
            profile[3 * r + 1] += dfy;
            profile[3 * r + 2] += dfz;
        }
    }
    memcpy(out, acc, sizeof(acc));
    return 0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Thread count of the following parallel loops (the CPU baseline's 1-core
   and all-core figures). */
int orc_set_threads(int n) {
#ifdef _OPENMP
    extern void omp_set_num_threads(int);
    omp_set_num_threads(n > 0 ? n : 1);
#endif
    return 0;
}

/* The reference's kernel part of a step, `steps` times, timed inside the
   library: collide f -> a (kernels.py:42-92 and its TRT / MRT variants),
   the post-collision copy a -> b (solver.py:348, f_pre = f_post.copy()),
   and the periodic push stream b -> f (kernels.py:94-115).  a and b are
   caller-owned work arrays of f's size.  Returns the wall seconds in *sec. */
int orc_kernel_only(int Q, int nz, int ny, int nx, double *f, double *a, double *b, const uint8_t *solid,
                    const int *ex, const int *ey, const int *ez, const double *w, const int *opp, int coll,
                    double omega, double trt_lambda, const double *M, const double *Minv, const double *s,
                    int trt_variant, int steps, double *sec) {
#ifdef _OPENMP
    extern double omp_get_wtime(void);
    const double t0 = omp_get_wtime();
#endif
    const long long n = (long long)Q * nz * ny * nx;
    for (int k = 0; k < steps; ++k) {
        orc_collide(Q, nz, ny, nx, f, a, solid, ex, ey, ez, w, opp, coll, omega, NULL, trt_lambda, M, Minv, s,
                    trt_variant);
#pragma omp parallel for schedule(static)
        for (long long j = 0; j < n; ++j) b[j] = a[j];
        orc_stream(Q, nz, ny, nx, b, f, ex, ey, ez);
    }
#ifdef _OPENMP
    *sec = omp_get_wtime() - t0;
#else
    *sec = 0.0;
#endif
    return 0;
}
