/*
 * bf_oracle_channel.c — CPU restatement of the differentiable channel and its
 * discrete adjoint.  TEST INFRASTRUCTURE ONLY (see bf_oracle.c).
 *
 * The reference has no counterpart: partial bounce-back, velocity/pressure
 * boundaries and the adjoint are new.  What is restated here is the published
 * algorithm each piece follows:
 *   - BGK collision with the equilibrium of collide.comp:589-606 (c_s^2 = 1/3);
 *   - partial bounce-back (Walsh, Burwinkle & Saar, Comput. Geosci. 35:1186,
 *     2009): f**_i = (1 - g) f*_i + g f_opp(i), with f* the BGK post-collision
 *     and f the pre-collision populations of the same cell;
 *   - Zou-He velocity inlet and pressure outlet (Zou & He, Phys. Fluids 9:1591,
 *     1997; D3Q19 transverse terms as in Hecht & Harting, J. Stat. Mech. 2010
 *     P01018): unknown f_i = f_opp(i) + s rho u_x wt_i - (c_iy N_y + c_iz N_z);
 *   - halfway bounce-back at the walls y = 0 and y = Y-1 (stream.comp:576-580
 *     with zero wall velocity); z periodic in 3-D;
 *   - objective J = mean over the outlet cells of u_x of the pre-collision
 *     populations at the final tick;
 *   - discrete adjoint: exact transpose of each stage (BC, collision + partial
 *     bounce-back, pull streaming), accumulated backwards over the stored
 *     trajectory.
 *
 * State convention: s^t = post-collision populations after tick t (stored per
 * cell, population-major); the pre-collision populations of tick t are pulled
 * from s^{t-1}; s^{-1} is the fluid at rest (f = w).  All populations are held
 * shifted, h = f - w, and rho = 1 + sum h, so FP32 sums carry no O(1) part
 * (the same FP32 accuracy device as the shifted 16-bit storage of the solver).
 * Every closure and the partial bounce-back keep their form in h because
 * w_i = w_opp(i) and sum_{c_x=0} w + 2 sum_{c_x=-s} w = 1.
 *
 * Parity pins (tests/test_oracle_channel.py): Zou-He reconstruction reproduces
 * the imposed moments; the adjoint gradient equals central finite differences
 * of this forward model; the dot-product identity per tick.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif
#define OCH_API __attribute__((visibility("default")))

typedef struct { int c[3]; double w; } cdir_t;

static const cdir_t CD2Q9[9] = {
    {{0, 0, 0}, 4.0 / 9}, {{1, 0, 0}, 1.0 / 9}, {{-1, 0, 0}, 1.0 / 9},
    {{0, 1, 0}, 1.0 / 9}, {{0, -1, 0}, 1.0 / 9}, {{1, 1, 0}, 1.0 / 36},
    {{-1, -1, 0}, 1.0 / 36}, {{1, -1, 0}, 1.0 / 36}, {{-1, 1, 0}, 1.0 / 36}};
static const cdir_t CD3Q19[19] = {
    {{0, 0, 0}, 1.0 / 3}, {{1, 0, 0}, 1.0 / 18}, {{-1, 0, 0}, 1.0 / 18},
    {{0, 1, 0}, 1.0 / 18}, {{0, -1, 0}, 1.0 / 18}, {{0, 0, 1}, 1.0 / 18},
    {{0, 0, -1}, 1.0 / 18}, {{1, 1, 0}, 1.0 / 36}, {{-1, -1, 0}, 1.0 / 36},
    {{1, -1, 0}, 1.0 / 36}, {{-1, 1, 0}, 1.0 / 36}, {{1, 0, 1}, 1.0 / 36},
    {{-1, 0, -1}, 1.0 / 36}, {{1, 0, -1}, 1.0 / 36}, {{-1, 0, 1}, 1.0 / 36},
    {{0, 1, 1}, 1.0 / 36}, {{0, -1, -1}, 1.0 / 36}, {{0, 1, -1}, 1.0 / 36},
    {{0, -1, 1}, 1.0 / 36}};

static int opp(int i) { return i == 0 ? 0 : ((i & 1) ? i + 1 : i - 1); }

typedef struct och {
    int X, Y, Z, q;
    long N;
    const cdir_t *T;
    REAL tau, u_in, rho_out;
    REAL *gamma;     /* N */
    REAL *prof;      /* Y: Poiseuille profile P(y), mean 1 over the fluid rows */
    REAL *state[2];  /* forward ping-pong */
    int cur;         /* state[cur] = s^{t-1} before a tick */
} och;

static long cid(const och *s, int x, int y, int z) { return ((long)z * s->Y + y) * s->X + x; }
static int is_wall(const och *s, int y) { return y == 0 || y == s->Y - 1; }

/* Unknown-population weight of the Zou-He closure: axis direction 2/3 (D2Q9)
 * or 1/3 (D3Q19), diagonals 1/6. */
static REAL zh_wt(const och *s, int i) {
    const int *c = s->T[i].c;
    int nz = (c[0] != 0) + (c[1] != 0) + (c[2] != 0);
    if (nz == 1) return s->q == 9 ? (REAL)2 / 3 : (REAL)1 / 3;
    return (REAL)1 / 6;
}

/* Pull the pre-collision populations of cell (x,y,z) from the post-collision
 * state `src`; slots pulled from outside the domain (x direction) are left 0
 * and filled by the boundary closure. */
static void pull(const och *s, const REAL *src, int x, int y, int z, REAL *f) {
    for (int i = 0; i < s->q; ++i) {
        const int *c = s->T[i].c;
        int sx = x - c[0], sy = y - c[1], sz = ((z - c[2]) % s->Z + s->Z) % s->Z;
        if (sx < 0 || sx >= s->X) { f[i] = 0; continue; }
        if (is_wall(s, sy)) { f[i] = src[(long)opp(i) * s->N + cid(s, x, y, z)]; continue; }
        f[i] = src[(long)i * s->N + cid(s, sx, sy, sz)];
    }
}

/* Boundary closure at the inlet (x = 0, sign +1) or outlet (x = X-1, sign -1).
 * Returns rho u_x; fills the unknown populations in f (shifted values).  S_out
 * receives the full sum S = 1 + S_h of the known populations (c_x = 0 once,
 * c_x = -sign twice). */
static REAL closure(const och *s, REAL *f, int sign, REAL ux_in, REAL *S_out, REAL *Ny, REAL *Nz) {
    REAL S = 0, ny = 0, nz = 0;
    for (int i = 0; i < s->q; ++i) {
        const int *c = s->T[i].c;
        if (c[0] == 0) { S += f[i]; ny += f[i] * c[1]; nz += f[i] * c[2]; }
        else if (c[0] == -sign) S += 2 * f[i];
    }
    ny *= (REAL)0.5; nz *= (REAL)0.5;
    REAL rux;
    if (sign > 0) rux = ((REAL)1 + S) * ux_in / ((REAL)1 - ux_in);   /* rho = S/(1-u) */
    else rux = S + ((REAL)1 - s->rho_out);                            /* rho u = rho_out (S/rho_out - 1) */
    for (int i = 0; i < s->q; ++i) {
        const int *c = s->T[i].c;
        if (c[0] != sign) continue;
        f[i] = f[opp(i)] + (REAL)sign * rux * zh_wt(s, i) - ((REAL)c[1] * ny + (REAL)c[2] * nz);
    }
    *S_out = (REAL)1 + S; *Ny = ny; *Nz = nz;
    return rux;
}

/* Collision + partial bounce-back of one cell; f (pre) -> out (post). Also
 * returns u and rho, and f* (BGK post-collision) in fstar when non-null. */
static void collide(const och *s, const REAL *f, REAL g, REAL *out, REAL *fstar, REAL *rho_o, REAL *u_o) {
    REAL delta = 0, m[3] = {0, 0, 0};
    for (int i = 0; i < s->q; ++i) {
        delta += f[i];
        m[0] += f[i] * s->T[i].c[0]; m[1] += f[i] * s->T[i].c[1]; m[2] += f[i] * s->T[i].c[2];
    }
    REAL rho = (REAL)1 + delta;
    REAL u[3] = {m[0] / rho, m[1] / rho, m[2] / rho};
    REAL uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    REAL fs[19];
    for (int i = 0; i < s->q; ++i) {
        const int *c = s->T[i].c;
        REAL cu = c[0] * u[0] + c[1] * u[1] + c[2] * u[2];
        /* f_eq - w = w (rho - 1) + w rho (3 cu + 9/2 cu^2 - 3/2 u^2) */
        REAL e = (REAL)s->T[i].w * (delta + rho * ((REAL)3 * cu + (REAL)4.5 * cu * cu - (REAL)1.5 * uu));
        fs[i] = f[i] - (f[i] - e) / s->tau;
    }
    for (int i = 0; i < s->q; ++i) out[i] = ((REAL)1 - g) * fs[i] + g * f[opp(i)];
    if (fstar) memcpy(fstar, fs, sizeof(REAL) * s->q);
    if (rho_o) *rho_o = rho;
    if (u_o) { u_o[0] = u[0]; u_o[1] = u[1]; u_o[2] = u[2]; }
}

/* Pre-collision populations of a fluid cell at tick t (pull + closures). */
static void precollision(const och *s, const REAL *src, int x, int y, int z, REAL *f) {
    pull(s, src, x, y, z, f);
    REAL S, ny, nz;
    if (x == 0) closure(s, f, +1, s->u_in * s->prof[y], &S, &ny, &nz);
    else if (x == s->X - 1) closure(s, f, -1, 0, &S, &ny, &nz);
}

static void tick(och *s, const REAL *src, REAL *dst) {
    /* rows are independent: parallel over (z, y) when built with OpenMP */
    const int rows = s->Z * (s->Y - 2);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; ++r)
    for (int x = 0; x < s->X; ++x) {
        REAL f[19], post[19];
        int z = r / (s->Y - 2), y = r % (s->Y - 2) + 1;
        long c = cid(s, x, y, z);
        precollision(s, src, x, y, z, f);
        collide(s, f, s->gamma[c], post, NULL, NULL, NULL);
        for (int i = 0; i < s->q; ++i) dst[(long)i * s->N + c] = post[i];
    }
}

static void rest_state(const och *s, REAL *dst) {
    memset(dst, 0, sizeof(REAL) * s->q * s->N);   /* f = w */
}

/* J = mean over outlet fluid cells of u_x of the pre-collision populations
 * pulled from `src` (= s^{T-1}). */
static REAL objective(const och *s, const REAL *src) {
    REAL f[19], acc = 0;
    long n = 0;
    for (int z = 0; z < s->Z; ++z)
    for (int y = 1; y < s->Y - 1; ++y) {
        REAL S, ny, nz;
        pull(s, src, s->X - 1, y, z, f);
        REAL rux = closure(s, f, -1, 0, &S, &ny, &nz);
        acc += rux / s->rho_out;
        ++n;
    }
    return acc / (REAL)n;
}

OCH_API och *och_create(int X, int Y, int Z, int vset, double tau, double u_in, double rho_out) {
    if (X < 2 || Y < 3 || Z < 1) return NULL;
    if (!(vset == 2 && Z == 1) && vset != 4) return NULL;
    och *s = (och *)calloc(1, sizeof *s);
    s->X = X; s->Y = Y; s->Z = Z; s->N = (long)X * Y * Z;
    s->q = vset == 2 ? 9 : 19;
    s->T = vset == 2 ? CD2Q9 : CD3Q19;
    s->tau = (REAL)tau; s->u_in = (REAL)u_in; s->rho_out = (REAL)rho_out;
    s->gamma = (REAL *)calloc(s->N, sizeof(REAL));
    s->prof = (REAL *)calloc(Y, sizeof(REAL));
    const double H = Y - 2.0;
    for (int y = 1; y < Y - 1; ++y) {
        double e = (y - 0.5) / H;
        s->prof[y] = (REAL)(6.0 * e * (1.0 - e));
    }
    s->state[0] = (REAL *)malloc(sizeof(REAL) * s->q * s->N);
    s->state[1] = (REAL *)malloc(sizeof(REAL) * s->q * s->N);
    return s;
}

OCH_API void och_destroy(och *s) {
    if (!s) return;
    free(s->gamma); free(s->prof); free(s->state[0]); free(s->state[1]); free(s);
}

OCH_API void och_set_porosity(och *s, const double *g) {
    for (long c = 0; c < s->N; ++c) s->gamma[c] = (REAL)g[c];
}

OCH_API void och_set_inlet(och *s, double u_in) { s->u_in = (REAL)u_in; }

OCH_API void och_profile(const och *s, double *out) {
    for (int y = 0; y < s->Y; ++y) out[y] = (double)s->prof[y];
}

/* Run T ticks from rest; returns J.  The last post-collision state is kept. */
OCH_API double och_forward(och *s, long T) {
    rest_state(s, s->state[0]);
    s->cur = 0;
    for (long t = 0; t < T; ++t) {
        tick(s, s->state[s->cur], s->state[s->cur ^ 1]);
        s->cur ^= 1;
    }
    return (double)objective(s, s->state[s->cur]);
}

/* Moments of the pre-collision populations after the last forward run:
 * out[4*cell + {0,1,2,3}] = (u_x, u_y, u_z, rho) for fluid cells, 0 for walls. */
OCH_API void och_moments(const och *s, double *out) {
    REAL f[19];
    memset(out, 0, sizeof(double) * 4 * s->N);
    for (int z = 0; z < s->Z; ++z)
    for (int y = 1; y < s->Y - 1; ++y)
    for (int x = 0; x < s->X; ++x) {
        precollision(s, s->state[s->cur], x, y, z, f);
        REAL rho = 1, m[3] = {0, 0, 0};
        for (int i = 0; i < s->q; ++i) {
            rho += f[i];
            m[0] += f[i] * s->T[i].c[0]; m[1] += f[i] * s->T[i].c[1]; m[2] += f[i] * s->T[i].c[2];
        }
        long c = cid(s, x, y, z);
        out[4 * c] = m[0] / rho; out[4 * c + 1] = m[1] / rho; out[4 * c + 2] = m[2] / rho; out[4 * c + 3] = rho;
    }
}

/* Pre-collision populations (unshifted) of a cell after closure, for the
 * closure tests: `src` is an arbitrary post-collision state (q*N, unshifted). */
OCH_API void och_precollision(const och *s, const double *src, int x, int y, int z, double *f) {
    REAL *tmp = (REAL *)malloc(sizeof(REAL) * s->q * s->N);
    for (long k = 0; k < (long)s->q * s->N; ++k) tmp[k] = (REAL)(src[k] - s->T[k / s->N].w);
    REAL ff[19];
    precollision(s, tmp, x, y, z, ff);
    for (int i = 0; i < s->q; ++i) f[i] = (double)ff[i] + s->T[i].w;
    free(tmp);
}

/* Adjoint of one tick at cell level: given nu = dJ/d s^t (q*N) and the state
 * s^{t-1} it was computed from, accumulate dJ/d s^{t-1} into nu_prev (which
 * must be zero on entry), dJ/d gamma into dg, and dJ/d u_in into *du. */
static void adjoint_tick(const och *s, const REAL *src, const REAL *nu, REAL *nu_prev, REAL *dg, REAL *du) {
    const int q = s->q;
    REAL f[19], fs[19], post[19], mu[19], mh[19];
    for (int z = 0; z < s->Z; ++z)
    for (int y = 1; y < s->Y - 1; ++y)
    for (int x = 0; x < s->X; ++x) {
        long c = cid(s, x, y, z);
        REAL S = 0, ny = 0, nz = 0, rux = 0;
        REAL uin = s->u_in * s->prof[y];
        pull(s, src, x, y, z, f);
        int sign = x == 0 ? 1 : (x == s->X - 1 ? -1 : 0);
        if (sign) rux = closure(s, f, sign, uin, &S, &ny, &nz);
        REAL rho, u[3], g = s->gamma[c];
        collide(s, f, g, post, fs, &rho, u);
        for (int i = 0; i < q; ++i) mu[i] = nu[(long)i * s->N + c];
        /* d/dgamma of post_i = f_opp - f*_i */
        REAL acc = 0;
        for (int i = 0; i < q; ++i) acc += mu[i] * (f[opp(i)] - fs[i]);
        dg[c] += acc;
        /* transpose of collision + partial bounce-back: O(q) via three sums */
        REAL uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        REAL A = 0, C = 0, B[3] = {0, 0, 0};
        for (int i = 0; i < q; ++i) {
            const int *ci = s->T[i].c;
            REAL cu = ci[0] * u[0] + ci[1] * u[1] + ci[2] * u[2];
            REAL wm = (REAL)s->T[i].w * mu[i];
            A += wm * ((REAL)1 - (REAL)4.5 * cu * cu + (REAL)1.5 * uu);
            C += wm;
            REAL b = wm * ((REAL)3 + (REAL)9 * cu);
            B[0] += b * ci[0]; B[1] += b * ci[1]; B[2] += b * ci[2];
        }
        REAL D[3] = {B[0] - 3 * C * u[0], B[1] - 3 * C * u[1], B[2] - 3 * C * u[2]};
        REAL a0 = ((REAL)1 - g) * ((REAL)1 - (REAL)1 / s->tau), a1 = ((REAL)1 - g) / s->tau;
        for (int j = 0; j < q; ++j) {
            const int *cj = s->T[j].c;
            mh[j] = a0 * mu[j] + g * mu[opp(j)] + a1 * (A + cj[0] * D[0] + cj[1] * D[1] + cj[2] * D[2]);
        }
        /* transpose of the closure: unknown slots feed the known ones */
        if (sign) {
            REAL gr = 0, gny = 0, gnz = 0;
            for (int i = 0; i < q; ++i) {
                const int *ci = s->T[i].c;
                if (ci[0] != sign) continue;
                mh[opp(i)] += mh[i];
                gr += mh[i] * (REAL)sign * zh_wt(s, i);
                gny -= mh[i] * ci[1];
                gnz -= mh[i] * ci[2];
                mh[i] = 0;
            }
            /* rux = S u/(1-u) (inlet) or S - rho_out (outlet) */
            REAL dS = sign > 0 ? gr * uin / ((REAL)1 - uin) : gr;
            if (sign > 0) du[0] += gr * S / (((REAL)1 - uin) * ((REAL)1 - uin)) * s->prof[y];
            for (int i = 0; i < q; ++i) {
                const int *ci = s->T[i].c;
                if (ci[0] == 0) mh[i] += dS + (REAL)0.5 * (gny * ci[1] + gnz * ci[2]);
                else if (ci[0] == -sign) mh[i] += 2 * dS;
            }
            (void)rux;
        }
        /* transpose of the pull: scatter to the slots of s^{t-1} */
        for (int i = 0; i < q; ++i) {
            const int *ci = s->T[i].c;
            int sx = x - ci[0], sy = y - ci[1], sz = ((z - ci[2]) % s->Z + s->Z) % s->Z;
            if (sx < 0 || sx >= s->X) continue;
            if (is_wall(s, sy)) nu_prev[(long)opp(i) * s->N + c] += mh[i];
            else nu_prev[(long)i * s->N + cid(s, sx, sy, sz)] += mh[i];
        }
    }
}

/* Seed dJ/d s^{T-1} from the objective. */
static void objective_adjoint(const och *s, REAL *nu) {
    long n = (long)(s->Y - 2) * s->Z;
    REAL k = (REAL)1 / ((REAL)n * s->rho_out);
    int x = s->X - 1;
    for (int z = 0; z < s->Z; ++z)
    for (int y = 1; y < s->Y - 1; ++y) {
        long c = cid(s, x, y, z);
        for (int i = 0; i < s->q; ++i) {
            const int *ci = s->T[i].c;
            REAL coef = ci[0] == 0 ? k : (ci[0] == 1 ? 2 * k : 0);
            if (coef == 0) continue;
            int sx = x - ci[0], sy = y - ci[1], sz = ((z - ci[2]) % s->Z + s->Z) % s->Z;
            if (is_wall(s, sy)) nu[(long)opp(i) * s->N + c] += coef;
            else nu[(long)i * s->N + cid(s, sx, sy, sz)] += coef;
        }
    }
}

/* Full gradient: forward storing every state, then the adjoint sweep.
 * Returns J; dgamma (N) and du (1) receive dJ/dgamma and dJ/du_in. */
OCH_API double och_gradient(och *s, long T, double *dgamma, double *du_out) {
    long qN = (long)s->q * s->N;
    REAL *traj = (REAL *)malloc(sizeof(REAL) * qN * (T + 1));  /* traj[t+1] = s^t, traj[0] = s^{-1} */
    rest_state(s, traj);
    for (long t = 0; t < T; ++t) tick(s, traj + t * qN, traj + (t + 1) * qN);
    double J = (double)objective(s, traj + T * qN);
    REAL *nu = (REAL *)calloc(qN, sizeof(REAL));
    REAL *nu_prev = (REAL *)calloc(qN, sizeof(REAL));
    REAL *dg = (REAL *)calloc(s->N, sizeof(REAL));
    REAL du = 0;
    objective_adjoint(s, nu);                    /* nu = dJ/d s^{T-1} */
    for (long t = T - 1; t >= 0; --t) {
        memset(nu_prev, 0, sizeof(REAL) * qN);
        adjoint_tick(s, traj + t * qN, nu, nu_prev, dg, &du);   /* tick t reads s^{t-1} = traj[t] */
        REAL *tmp = nu; nu = nu_prev; nu_prev = tmp;
    }
    for (long c = 0; c < s->N; ++c) dgamma[c] = (double)dg[c];
    *du_out = (double)du;
    /* keep the final state for och_moments */
    memcpy(s->state[0], traj + T * qN, sizeof(REAL) * qN);
    s->cur = 0;
    free(traj); free(nu); free(nu_prev); free(dg);
    return J;
}

/* The GPU's checkpointed schedule: the forward keeps the state every K ticks;
 * the adjoint recomputes each segment from its checkpoint.  half_states = 1
 * keeps the recomputed states j >= 1 of a segment as IEEE binary16 (round to
 * nearest even of the shifted value, through FP32), and the recompute
 * continues from the rounded state, as the device does with 16-bit segment
 * storage; half_states = 0 is exact and equals och_gradient. */
float orc_half_rne(float v);

OCH_API double och_gradient_checkpointed(och *s, long T, long K, int half_states, double *dgamma, double *du_out) {
    const long qN = (long)s->q * s->N;
    const long nseg = (T + K - 1) / K;
    REAL *ck = (REAL *)malloc(sizeof(REAL) * qN * nseg);
    REAL *a = (REAL *)malloc(sizeof(REAL) * qN), *b = (REAL *)malloc(sizeof(REAL) * qN);
    REAL *seg = (REAL *)malloc(sizeof(REAL) * qN * K);
    rest_state(s, a);
    memcpy(ck, a, sizeof(REAL) * qN);
    for (long t = 0; t < T; ++t) {
        tick(s, a, b);
        REAL *tmp = a; a = b; b = tmp;
        if ((t + 1) % K == 0 && t + 1 < T) memcpy(ck + ((t + 1) / K) * qN, a, sizeof(REAL) * qN);
    }
    double J = (double)objective(s, a);
    memcpy(s->state[0], a, sizeof(REAL) * qN);
    s->cur = 0;
    REAL *nu = (REAL *)calloc(qN, sizeof(REAL));
    REAL *nu_prev = (REAL *)calloc(qN, sizeof(REAL));
    REAL *dg = (REAL *)calloc(s->N, sizeof(REAL));
    REAL du = 0;
    objective_adjoint(s, nu);
    for (long k = nseg - 1; k >= 0; --k) {
        const long len = (T - k * K) < K ? (T - k * K) : K;
        memcpy(seg, ck + k * qN, sizeof(REAL) * qN);
        for (long j = 1; j < len; ++j) {
            tick(s, seg + (j - 1) * qN, seg + j * qN);
            if (half_states)
                for (long e = 0; e < qN; ++e) seg[j * qN + e] = (REAL)orc_half_rne((float)seg[j * qN + e]);
        }
        for (long j = len - 1; j >= 0; --j) {
            memset(nu_prev, 0, sizeof(REAL) * qN);
            adjoint_tick(s, seg + j * qN, nu, nu_prev, dg, &du);
            REAL *tmp = nu; nu = nu_prev; nu_prev = tmp;
        }
    }
    for (long c = 0; c < s->N; ++c) dgamma[c] = (double)dg[c];
    *du_out = (double)du;
    free(ck); free(a); free(b); free(seg); free(nu); free(nu_prev); free(dg);
    return J;
}

/* One tick and its adjoint on arbitrary inputs (shifted values), for the
 * dot-product test: out = tick(src); nu_prev = (d tick / d src)^T nu. */
OCH_API void och_tick_pair(och *s, const double *src, const double *nu, double *out, double *nu_prev) {
    long qN = (long)s->q * s->N;
    REAL *a = (REAL *)calloc(qN, sizeof(REAL)), *b = (REAL *)calloc(qN, sizeof(REAL));
    REAL *n = (REAL *)calloc(qN, sizeof(REAL)), *np = (REAL *)calloc(qN, sizeof(REAL));
    REAL *dg = (REAL *)calloc(s->N, sizeof(REAL));
    REAL du = 0;
    for (long k = 0; k < qN; ++k) { a[k] = (REAL)src[k]; n[k] = (REAL)nu[k]; }
    tick(s, a, b);
    adjoint_tick(s, a, n, np, dg, &du);
    for (long k = 0; k < qN; ++k) { out[k] = (double)b[k]; nu_prev[k] = (double)np[k]; }
    free(a); free(b); free(n); free(np); free(dg);
}
