/*
 * CPU FP64 oracle for the D2Q9 lattice-Boltzmann step, C restatement.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Used by tests/ for long runs that
 * the numpy oracle (oracle/lbm_oracle.py) would take minutes to finish, and by
 * bench.py's cpu_baseline leg.  Nothing in the product links it.
 *
 * It performs the same IEEE operations in the same order as oracle/lbm_oracle.py
 * (which restates src/lattice_boltzmann_method.py:93-228 and
 * src/boundary_conditions.py:78-348 of the reference), so the two are
 * bit-identical; tests/test_oracle.py
 * (test_c_oracle_equals_numpy_oracle_on_all_composites) checks that.  Build with
 * -ffp-contract=off and without -ffast-math: a fused multiply-add would change
 * the rounding.
 *
 * Layout: AoS f[x][y][i], C order, lx*ly*9 doubles, as in the reference.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static const int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int CY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
static const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
static const double W[9] = {4.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9, 1.0 / 9,
                            1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

/* op kinds, in the order the composite applies them */
enum { OP_REFLECT = 1, OP_INLET = 2, OP_OUTLET = 3, OP_PRESSURE = 4 };

typedef struct {
    int kind;
    int dirs;          /* REFLECT: source directions i (bit i); PRESSURE: D1 bits */
    int dirs2;         /* PRESSURE: D2 bits */
    long ncells;       /* REFLECT: number of cells */
    const int64_t *cells; /* REFLECT: linear indices x*ly + y */
    double k[9];       /* REFLECT: per-direction constant; INLET: feq */
    double rho_in, rho_out; /* PRESSURE */
} oracle_op;

static int g_square_pow = 0;

static inline double sq(double v) { return g_square_pow ? pow(v, 2.0) : v * v; }

/* feq for one cell, same expression tree as equilibrium() in lbm_oracle.py */
static inline void feq_cell(double rho, double ux, double uy, double *out) {
    double usq = sq(sqrt(ux * ux + uy * uy));
    for (int i = 0; i < 9; i++) {
        double cu = ux * (double)CX[i] + uy * (double)CY[i];
        out[i] = (W[i] * rho) * (((1.0 + 3.0 * cu) + 4.5 * sq(cu)) - 1.5 * usq);
    }
}

static inline double rho_cell(const double *a) {
    return (((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]))) + a[8];
}

static inline void vel_cell(double rho, const double *a, double *ux, double *uy) {
    double jx = ((a[1] + a[5]) + a[8]) - ((a[3] + a[6]) + a[7]);
    double jy = ((a[2] + a[5]) + a[6]) - ((a[4] + a[7]) + a[8]);
    if (rho != 0.0) {
        *ux = jx / rho;
        *uy = jy / rho;
    } else {
        *ux = 0.0;
        *uy = 0.0;
    }
}

void oracle_set_square_pow(int on) { g_square_pow = on; }

void oracle_equilibrium(const double *rho, const double *u, double *feq, long n) {
    for (long c = 0; c < n; c++) feq_cell(rho[c], u[2 * c], u[2 * c + 1], feq + 9 * c);
}

/*
 * Run nsteps reference steps in place on (f, rho, u).
 * probe >= 0: record (rho, ux, uy) of that cell after every step into trace.
 * Returns 0, or -1 on allocation failure.
 */
int oracle_run(int lx, int ly, double *f, double *rho, double *u, double omega, long nsteps,
               const oracle_op *ops, int nops, long probe, double *trace) {
    const long n = (long)lx * ly;
    double *pre = (double *)malloc(sizeof(double) * 9 * n);
    double *post = (double *)malloc(sizeof(double) * 9 * n);
    if (!pre || !post) {
        free(pre);
        free(post);
        return -1;
    }
    for (long s = 0; s < nsteps; s++) {
        /* collide with the step-input moments */
#pragma omp parallel for schedule(static)
        for (long c = 0; c < n; c++) {
            double e[9];
            feq_cell(rho[c], u[2 * c], u[2 * c + 1], e);
            for (int i = 0; i < 9; i++) pre[9 * c + i] = f[9 * c + i] + (e[i] - f[9 * c + i]) * omega;
        }
        /* pre-streaming pressure fix (columns 0 and lx-1) */
        for (int k = 0; k < nops; k++) {
            const oracle_op *op = &ops[k];
            if (op->kind != OP_PRESSURE) continue;
            const int xa = (lx - 2 + lx) % lx, xb = 1 % lx;
            for (int y = 0; y < ly; y++) {
                long ca = (long)xa * ly + y;
                double ein[9], ea[9];
                feq_cell(op->rho_in, u[2 * ca], u[2 * ca + 1], ein);
                feq_cell(rho[ca], u[2 * ca], u[2 * ca + 1], ea);
                for (int i = 0; i < 9; i++)
                    if (op->dirs >> i & 1)
                        pre[9 * ((long)0 * ly + y) + i] = ein[i] + (pre[9 * ca + i] - ea[i]);
            }
            for (int y = 0; y < ly; y++) {
                long cb = (long)xb * ly + y;
                double eout[9], eb[9];
                feq_cell(op->rho_out, u[2 * cb], u[2 * cb + 1], eout);
                feq_cell(rho[cb], u[2 * cb], u[2 * cb + 1], eb);
                for (int i = 0; i < 9; i++)
                    if (op->dirs2 >> i & 1)
                        pre[9 * ((long)(lx - 1) * ly + y) + i] = eout[i] + (pre[9 * cb + i] - eb[i]);
            }
        }
        /* periodic pull streaming */
#pragma omp parallel for schedule(static)
        for (int x = 0; x < lx; x++) {
            for (int y = 0; y < ly; y++) {
                long c = (long)x * ly + y;
                for (int i = 0; i < 9; i++) {
                    int sx = (x - CX[i] + lx) % lx, sy = (y - CY[i] + ly) % ly;
                    post[9 * c + i] = pre[9 * ((long)sx * ly + sy) + i];
                }
            }
        }
        /* post-streaming boundary ops, last writer wins */
        for (int k = 0; k < nops; k++) {
            const oracle_op *op = &ops[k];
            if (op->kind == OP_REFLECT) {
                for (int i = 0; i < 9; i++) {
                    if (!(op->dirs >> i & 1)) continue;
                    for (long j = 0; j < op->ncells; j++) {
                        long c = op->cells[j];
                        post[9 * c + OPP[i]] = pre[9 * c + i] - op->k[i];
                    }
                }
            } else if (op->kind == OP_INLET) {
                for (int y = 0; y < ly; y++)
                    for (int i = 0; i < 9; i++) post[9 * (long)y + i] = op->k[i];
            } else if (op->kind == OP_OUTLET) {
                const int xs = (lx - 2 + lx) % lx;
                for (int y = 0; y < ly; y++) {
                    long cd = (long)(lx - 1) * ly + y, cs = (long)xs * ly + y;
                    post[9 * cd + 3] = f[9 * cs + 3];
                    post[9 * cd + 6] = f[9 * cs + 6];
                    post[9 * cd + 7] = f[9 * cs + 7];
                }
            }
        }
        /* moments of the post-boundary populations */
#pragma omp parallel for schedule(static)
        for (long c = 0; c < n; c++) {
            double r = rho_cell(post + 9 * c);
            rho[c] = r;
            vel_cell(r, post + 9 * c, &u[2 * c], &u[2 * c + 1]);
        }
        memcpy(f, post, sizeof(double) * 9 * n);
        if (probe >= 0 && trace) {
            trace[3 * s] = rho[probe];
            trace[3 * s + 1] = u[2 * probe];
            trace[3 * s + 2] = u[2 * probe + 1];
        }
    }
    free(pre);
    free(post);
    return 0;
}

/* |u| as numpy's 1-D norm rounds it through an FMA-based BLAS dot:
 * sqrt(fma(uy, uy, ux*ux)).  The reference's MPI probe traces were written
 * that way (tests/test_oracle.py checks it against them). */
void oracle_norm_fma(const double *ux, const double *uy, double *out, long n) {
    for (long i = 0; i < n; i++) out[i] = sqrt(fma(uy[i], uy[i], ux[i] * ux[i]));
}

/* OpenMP threads a parallel region of oracle_run uses. */
int oracle_threads(void) { return omp_get_max_threads(); }
