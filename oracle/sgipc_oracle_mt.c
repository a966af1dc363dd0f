/*
 * sgipc_oracle_mt.c -- multi-threaded variants of the oracle's heavy stages,
 * for the CPU baseline timed on all host cores.  TEST INFRASTRUCTURE ONLY
 * (see sgipc_oracle.h).  or_set_threads(t > 1) routes or_fem_assemble,
 * or_convert, or_mas_build and or_pcg here; t = 1 runs the serial code.
 *
 * Same algorithms as the serial restatement (S/femEnergy.cu:1719-1857,
 * S/linear_system/utils/converter.cu:23-196, S/MASPreconditioner.cu:
 * 1772-2262, S/linear_system/solver/pcg_solver.cu:174-267).  Results:
 *   - FEM triplets: identical; the vertex gradient is reduced per thread
 *     then summed in thread order (round-off differences only);
 *   - convert: identical (each key summed in ascending triplet order);
 *   - MAS tables, bank matrices and inverses: identical (every bank
 *     accumulates its sources in the serial order);
 *   - MAS apply: identical (nodes sum their vertices in ascending order);
 *   - PCG: the SpMV sums a row's blocks in a different order and the dots
 *     are blocked per thread, so iterates agree to round-off.
 */
#include "sgipc_oracle_internal.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

int or_threads = 1;

/* ---- a persistent pool: run f(ctx, tid, lo, hi) over [0, n) in T chunks -- */
typedef void (*or_task)(void* ctx, int tid, long long lo, long long hi);

static struct {
    int nthreads;                 /* workers + the caller */
    pthread_t* tid;
    pthread_barrier_t start, done;
    or_task task;
    void* ctx;
    long long n;
    int quit;
} pool;

static void chunk_of(long long n, int T, int t, long long* lo, long long* hi) {
    *lo = n * t / T;
    *hi = n * (t + 1) / T;
}

static void* worker(void* arg) {
    int t = (int)(size_t)arg;
    for (;;) {
        pthread_barrier_wait(&pool.start);
        if (pool.quit) break;
        long long lo, hi;
        chunk_of(pool.n, pool.nthreads, t, &lo, &hi);
        if (lo < hi) pool.task(pool.ctx, t, lo, hi);
        pthread_barrier_wait(&pool.done);
    }
    return NULL;
}

static void pool_stop(void) {
    if (pool.nthreads <= 1) return;
    pool.quit = 1;
    pthread_barrier_wait(&pool.start);
    for (int t = 1; t < pool.nthreads; ++t) pthread_join(pool.tid[t], NULL);
    pthread_barrier_destroy(&pool.start);
    pthread_barrier_destroy(&pool.done);
    free(pool.tid);
    memset(&pool, 0, sizeof(pool));
}

static void pool_start(int T) {
    pool_stop();
    pool.nthreads = T;
    pool.quit = 0;
    if (T <= 1) return;
    pool.tid = (pthread_t*)calloc((size_t)T, sizeof(pthread_t));
    pthread_barrier_init(&pool.start, NULL, (unsigned)T);
    pthread_barrier_init(&pool.done, NULL, (unsigned)T);
    for (int t = 1; t < T; ++t) pthread_create(&pool.tid[t], NULL, worker, (void*)(size_t)t);
}

static void par_for(long long n, or_task f, void* ctx) {
    if (n <= 0) return;
    if (pool.nthreads <= 1) {
        f(ctx, 0, 0, n);
        return;
    }
    pool.task = f;
    pool.ctx = ctx;
    pool.n = n;
    pthread_barrier_wait(&pool.start);
    long long lo, hi;
    chunk_of(n, pool.nthreads, 0, &lo, &hi);
    if (lo < hi) f(ctx, 0, lo, hi);
    pthread_barrier_wait(&pool.done);
}

void or_set_threads(int t) {
    if (t < 1) t = 1;
    if (t > 256) t = 256;
    or_threads = t;
    if (t != pool.nthreads) pool_start(t);
}
int or_get_threads(void) { return or_threads; }

/* ---- FEM assembly --------------------------------------------------------- */
typedef struct {
    int n;
    const double *x, *DmInv, *vol, *mu, *lam;
    const uint32_t* tets;
    double dt;
    const int* btype;
    int off;
    double *grad, *vals, **tgrad;
    int *rows, *cols;
} fem_ctx;

static void fem_task(void* c_, int tid, long long lo, long long hi) {
    fem_ctx* c = (fem_ctx*)c_;
    double xt[12], g[12], H[144];
    double* tg = c->tgrad[tid];
    memset(tg, 0, sizeof(double) * 3 * (size_t)c->n);
    for (long long t = lo; t < hi; ++t) {
        const uint32_t* v = c->tets + 4 * t;
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) xt[3 * a + k] = c->x[3 * v[a] + k];
        or_snk1_tet(xt, c->DmInv + 9 * t, c->vol[t], c->mu[t], c->lam[t], c->dt, g, H);
        for (int a = 0; a < 4; ++a)
            for (int k = 0; k < 3; ++k) tg[3 * v[a] + k] += g[3 * a + k];
        int kk = 0;
        for (int ii = 0; ii < 4; ++ii)
            for (int jj = ii; jj < 4; ++jj, ++kk) {
                int r = (int)v[ii] + c->off, cc = (int)v[jj] + c->off, tr = 0;
                if (r > cc) { int s = r; r = cc; cc = s; tr = 1; }
                int zero = c->btype && (c->btype[r - c->off] != 0 || c->btype[cc - c->off] != 0);
                double* B = c->vals + 9 * (10 * (size_t)t + kk);
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) {
                        double h = tr ? H[12 * (3 * jj + p) + 3 * ii + q] : H[12 * (3 * ii + p) + 3 * jj + q];
                        B[p + 3 * q] = zero ? 0.0 : h;
                    }
                c->rows[10 * (size_t)t + kk] = r;
                c->cols[10 * (size_t)t + kk] = cc;
            }
    }
}

static void fem_reduce_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    fem_ctx* c = (fem_ctx*)c_;
    for (long long i = lo; i < hi; ++i) {
        double s = 0.0;
        for (int t = 0; t < pool.nthreads; ++t) s += c->tgrad[t][i];
        c->grad[i] += s;
    }
}

void or_fem_assemble_mt(int n, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                        const double* mu, const double* lam, double dt, const int* btype, int off, double* grad,
                        double* vals, int* rows, int* cols) {
    int nt = pool.nthreads > 0 ? pool.nthreads : 1;
    double** tg = (double**)malloc(sizeof(double*) * (size_t)nt);
    for (int t = 0; t < nt; ++t) tg[t] = (double*)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    fem_ctx c = {n, x, DmInv, vol, mu, lam, tets, dt, btype, off, grad, vals, tg, rows, cols};
    /* a thread whose chunk is empty leaves its buffer untouched: clear all */
    for (int t = 0; t < nt; ++t) memset(tg[t], 0, sizeof(double) * 3 * (size_t)n);
    par_for(T, fem_task, &c);
    par_for(3LL * n, fem_reduce_task, &c);
    for (int t = 0; t < nt; ++t) free(tg[t]);
    free(tg);
}

/* ---- triplet -> BCOO ------------------------------------------------------ */
typedef struct { uint64_t key; int t; } kt_t;
static int kt_cmp(const void* a, const void* b) {
    const kt_t *x = (const kt_t*)a, *y = (const kt_t*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return (x->t > y->t) - (x->t < y->t);
}

typedef struct {
    int nt, nbk;
    long long nrows;
    const int *rows, *cols;
    const double* vals;
    long long* cnt;   /* T x nbk, then offsets */
    kt_t* e;
    long long* bstart; /* nbk + 1 */
    int* ucount;       /* unique keys per bucket */
    long long* ustart; /* nbk + 1 */
    int *orow, *ocol;
    double* oval;
    int* rmax;
} conv_ctx;

static int bucket_of(const conv_ctx* c, int r) { return (int)((long long)r * c->nbk / c->nrows); }

static void conv_max_task(void* c_, int tid, long long lo, long long hi) {
    conv_ctx* c = (conv_ctx*)c_;
    int m = 0;
    for (long long i = lo; i < hi; ++i) if (c->rows[i] > m) m = c->rows[i];
    c->rmax[tid] = m;
}
static void conv_count_task(void* c_, int tid, long long lo, long long hi) {
    conv_ctx* c = (conv_ctx*)c_;
    long long* cn = c->cnt + (size_t)tid * c->nbk;
    for (long long i = lo; i < hi; ++i) cn[bucket_of(c, c->rows[i])]++;
}
static void conv_scatter_task(void* c_, int tid, long long lo, long long hi) {
    conv_ctx* c = (conv_ctx*)c_;
    long long* pos = c->cnt + (size_t)tid * c->nbk;
    for (long long i = lo; i < hi; ++i) {
        int b = bucket_of(c, c->rows[i]);
        kt_t* d = c->e + pos[b]++;
        d->key = ((uint64_t)(uint32_t)c->rows[i] << 32) + (uint64_t)(uint32_t)c->cols[i];
        d->t = (int)i;
    }
}
static void conv_sort_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    conv_ctx* c = (conv_ctx*)c_;
    for (long long b = lo; b < hi; ++b) {
        kt_t* e = c->e + c->bstart[b];
        long long m = c->bstart[b + 1] - c->bstart[b];
        if (m > 1) qsort(e, (size_t)m, sizeof(kt_t), kt_cmp);
        int u = 0;
        for (long long i = 0; i < m; ++i) u += (i == 0 || e[i].key != e[i - 1].key);
        c->ucount[b] = u;
    }
}
static void conv_sum_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    conv_ctx* c = (conv_ctx*)c_;
    for (long long b = lo; b < hi; ++b) {
        kt_t* e = c->e + c->bstart[b];
        long long m = c->bstart[b + 1] - c->bstart[b];
        long long K = c->ustart[b] - 1;
        for (long long i = 0; i < m; ++i) {
            const double* v = c->vals + 9 * (size_t)e[i].t;
            if (i == 0 || e[i].key != e[i - 1].key) {
                ++K;
                c->orow[K] = (int)(e[i].key >> 32);
                c->ocol[K] = (int)(e[i].key & 0xffffffffu);
                for (int k = 0; k < 9; ++k) c->oval[9 * (size_t)K + k] = v[k];
            } else {
                for (int k = 0; k < 9; ++k) c->oval[9 * (size_t)K + k] += v[k];
            }
        }
    }
}

/* rows bucketed by contiguous ranges (stable scatter in triplet order),
 * each bucket sorted by (key, triplet) and summed: the serial result */
int or_convert_mt(int nt, const int* rows, const int* cols, const double* vals, int* orow, int* ocol, double* oval) {
    int T = pool.nthreads > 0 ? pool.nthreads : 1;
    conv_ctx c;
    memset(&c, 0, sizeof(c));
    c.nt = nt; c.rows = rows; c.cols = cols; c.vals = vals; c.orow = orow; c.ocol = ocol; c.oval = oval;
    c.rmax = (int*)calloc((size_t)T, sizeof(int));
    par_for(nt, conv_max_task, &c);
    int rmax = 0;
    for (int t = 0; t < T; ++t) if (c.rmax[t] > rmax) rmax = c.rmax[t];
    c.nrows = (long long)rmax + 1;
    c.nbk = 8 * T;
    c.cnt = (long long*)calloc((size_t)T * c.nbk, sizeof(long long));
    par_for(nt, conv_count_task, &c);
    /* offsets: bucket-major, chunk order inside a bucket (stable) */
    c.bstart = (long long*)malloc(sizeof(long long) * ((size_t)c.nbk + 1));
    long long acc = 0;
    for (int b = 0; b < c.nbk; ++b) {
        c.bstart[b] = acc;
        for (int t = 0; t < T; ++t) {
            long long v = c.cnt[(size_t)t * c.nbk + b];
            c.cnt[(size_t)t * c.nbk + b] = acc;
            acc += v;
        }
    }
    c.bstart[c.nbk] = acc;
    c.e = (kt_t*)malloc(sizeof(kt_t) * (size_t)nt);
    par_for(nt, conv_scatter_task, &c);
    c.ucount = (int*)calloc((size_t)c.nbk, sizeof(int));
    par_for(c.nbk, conv_sort_task, &c);
    c.ustart = (long long*)malloc(sizeof(long long) * ((size_t)c.nbk + 1));
    acc = 0;
    for (int b = 0; b < c.nbk; ++b) { c.ustart[b] = acc; acc += c.ucount[b]; }
    c.ustart[c.nbk] = acc;
    par_for(c.nbk, conv_sum_task, &c);
    free(c.rmax); free(c.cnt); free(c.bstart); free(c.e); free(c.ucount); free(c.ustart);
    return (int)acc;
}

/* ---- MAS Galerkin projection + inverses ------------------------------------ */
typedef struct {
    or_mas* m;
    int K, offset, nbanks, T;
    const int *rows, *cols;
    const double* vals;
    int* bank_lk;       /* L x K: target bank of source k at level l, or -1 */
    long long* cnt;     /* T x T: sources of chunk t for owner thread u */
    unsigned* src;      /* per owner: (l << 29) | k, serial order */
    long long* ostart;  /* T + 1 */
} proj_ctx;

static int owner_of(const proj_ctx* c, int bank) { return (int)((long long)bank * c->T / c->nbanks); }

static void proj_bank_task(void* c_, int tid, long long lo, long long hi) {
    proj_ctx* c = (proj_ctx*)c_;
    or_mas* m = c->m;
    const int B = m->B;
    long long* cn = c->cnt + (size_t)tid * c->T;
    for (int l = 0; l < m->L; ++l)
        for (long long k = lo; k < hi; ++k) {
            int r = c->rows[k] - c->offset, cc = c->cols[k] - c->offset, bank = -1;
            if (r >= 0 && cc >= 0 && r < m->n && cc < m->n) {
                int I = or__node_of(m, r, l), J = or__node_of(m, cc, l);
                if (I / B == J / B) bank = I / B;
            }
            c->bank_lk[(size_t)l * c->K + k] = bank;
            if (bank >= 0) cn[owner_of(c, bank)]++;
        }
}

static void proj_scatter_task(void* c_, int tid, long long lo, long long hi) {
    proj_ctx* c = (proj_ctx*)c_;
    long long* pos = c->cnt + (size_t)tid * c->T;
    for (int l = 0; l < c->m->L; ++l)
        for (long long k = lo; k < hi; ++k) {
            int bank = c->bank_lk[(size_t)l * c->K + k];
            if (bank >= 0) c->src[pos[owner_of(c, bank)]++] = ((unsigned)l << 29) | (unsigned)k;
        }
}

/* owner thread u accumulates the sources of its banks; a bank belongs to
 * one level, and its sources arrive in ascending k as in the serial loop */
static void proj_accum_task(void* c_, int tid, long long lo, long long hi) {
    (void)lo; (void)hi;
    proj_ctx* c = (proj_ctx*)c_;
    or_mas* m = c->m;
    const int B = m->B, D = 3 * B;
    const size_t bs = (size_t)D * D;
    for (long long s = c->ostart[tid]; s < c->ostart[tid + 1]; ++s) {
        int l = (int)(c->src[s] >> 29), k = (int)(c->src[s] & ((1u << 29) - 1u));
        int r = c->rows[k] - c->offset, cc = c->cols[k] - c->offset;
        const double* H = c->vals + 9 * (size_t)k;
        for (int tr = 0; tr < 2; ++tr) {
            if (tr == 1 && r == cc) break;
            int I = or__node_of(m, tr ? cc : r, l), J = or__node_of(m, tr ? r : cc, l);
            double* A = m->A + bs * (size_t)(I / B);
            int bi = (I % B) * 3, bj = (J % B) * 3;
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) {
                    double h = tr ? H[q + 3 * p] : H[p + 3 * q];
                    A[(size_t)(bi + p) * D + bj + q] += h;
                }
        }
    }
}

static void invert_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    or_mas* m = (or_mas*)c_;
    int D = 3 * m->B;
    double* S = (double*)malloc(sizeof(double) * (size_t)D * D);
    double* colm = (double*)malloc(sizeof(double) * (size_t)D);
    for (long long b = lo; b < hi; ++b) or__invert_bank(m, (int)b, S, colm);
    free(S); free(colm);
}

static void one_owner_task(void* c_, int tid, long long lo, long long hi) {
    (void)lo; (void)hi;
    proj_accum_task(c_, tid, 0, 1);
}

void or_mas_project_invert_mt(or_mas* m, int K, const int* rows, const int* cols, const double* vals, int offset) {
    int T = pool.nthreads > 0 ? pool.nthreads : 1;
    proj_ctx c;
    memset(&c, 0, sizeof(c));
    c.m = m; c.K = K; c.offset = offset; c.rows = rows; c.cols = cols; c.vals = vals; c.T = T;
    c.nbanks = m->total / m->B;
    if (c.nbanks > 0 && K > 0) {
        c.bank_lk = (int*)malloc(sizeof(int) * (size_t)m->L * K);
        c.cnt = (long long*)calloc((size_t)T * T, sizeof(long long));
        par_for(K, proj_bank_task, &c);
        /* owner-major, chunk order within an owner */
        c.ostart = (long long*)malloc(sizeof(long long) * ((size_t)T + 1));
        long long acc = 0;
        for (int u = 0; u < T; ++u) {
            c.ostart[u] = acc;
            for (int t = 0; t < T; ++t) {
                long long v = c.cnt[(size_t)t * T + u];
                c.cnt[(size_t)t * T + u] = acc;
                acc += v;
            }
        }
        c.ostart[T] = acc;
        c.src = (unsigned*)malloc(sizeof(unsigned) * (size_t)(acc > 0 ? acc : 1));
        par_for(K, proj_scatter_task, &c);
        par_for(T, one_owner_task, &c);  /* one owner per thread */
        free(c.bank_lk); free(c.cnt); free(c.ostart); free(c.src);
    }
    par_for(c.nbanks, invert_task, m);
}

/* ---- PCG ------------------------------------------------------------------ */
typedef struct {
    int K, nb, N;
    const int *rows, *cols;
    const double* vals;
    int* rptr;    /* nb + 1: upper blocks of row i (BCOO is sorted by row) */
    int* cptr;    /* nb + 1: strictly-upper blocks with column j */
    int* cidx;    /* their block indices, ascending */
    long long* cnt;
    const double* xin;
    double* yout;
    double a, b;
    double* part;  /* per thread */
    const double *u, *v;
    double alpha;
    double *x, *r, *p, *z, *Ap;
    or_mas* mas;
    int mas_fp32;
    double *R, *Z;
    float *Rf, *Zf;
    const double* rin;
    double* zout;
    int level;
    const double* Dinv;
    int precond;
} pcg_ctx;

static void rowptr_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    pcg_ctx* c = (pcg_ctx*)c_;
    /* rptr[i] = first k with rows[k] >= i (binary search) */
    for (long long i = lo; i < hi; ++i) {
        int a = 0, b = c->K;
        while (a < b) {
            int mid = a + (b - a) / 2;
            if (c->rows[mid] < i) a = mid + 1; else b = mid;
        }
        c->rptr[i] = a;
    }
}
static void ccount_task(void* c_, int tid, long long lo, long long hi) {
    pcg_ctx* c = (pcg_ctx*)c_;
    int* cn = (int*)(c->cnt) + 0;
    (void)cn;
    long long* mine = c->cnt + (size_t)tid * (c->nb + 1);
    for (long long k = lo; k < hi; ++k)
        if (c->rows[k] != c->cols[k]) mine[c->cols[k]]++;
}
static void cfill_task(void* c_, int tid, long long lo, long long hi) {
    pcg_ctx* c = (pcg_ctx*)c_;
    long long* pos = c->cnt + (size_t)tid * (c->nb + 1);
    for (long long k = lo; k < hi; ++k)
        if (c->rows[k] != c->cols[k]) c->cidx[pos[c->cols[k]]++] = (int)k;
}

/* y = a A x + b y, row i: its upper blocks, then the transposed blocks of
 * column i */
static void spmv_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    pcg_ctx* c = (pcg_ctx*)c_;
    const double* x = c->xin;
    for (long long i = lo; i < hi; ++i) {
        double s[3] = {0, 0, 0};
        for (int k = c->rptr[i]; k < c->rptr[i + 1]; ++k) {
            const double* M = c->vals + 9 * (size_t)k;
            int j = c->cols[k];
            for (int p = 0; p < 3; ++p)
                s[p] += M[p] * x[3 * j] + M[p + 3] * x[3 * j + 1] + M[p + 6] * x[3 * j + 2];
        }
        for (int e = c->cptr[i]; e < c->cptr[i + 1]; ++e) {
            int k = c->cidx[e], j = c->rows[k];
            const double* M = c->vals + 9 * (size_t)k;
            for (int q = 0; q < 3; ++q)
                s[q] += M[3 * q] * x[3 * j] + M[1 + 3 * q] * x[3 * j + 1] + M[2 + 3 * q] * x[3 * j + 2];
        }
        for (int p = 0; p < 3; ++p) {
            double* y = c->yout + 3 * i + p;
            *y = (c->b != 0.0 ? c->b * *y : 0.0) + c->a * s[p];
        }
    }
}
static void spmv_mt(pcg_ctx* c, double a, const double* x, double b, double* y) {
    c->a = a; c->b = b; c->xin = x; c->yout = y;
    par_for(c->nb, spmv_task, c);
}

static void dot_task(void* c_, int tid, long long lo, long long hi) {
    pcg_ctx* c = (pcg_ctx*)c_;
    double s = 0.0;
    for (long long i = lo; i < hi; ++i) s += c->u[i] * c->v[i];
    c->part[tid] = s;
}
static double dot_mt(pcg_ctx* c, const double* u, const double* v) {
    int T = pool.nthreads > 0 ? pool.nthreads : 1;
    for (int t = 0; t < T; ++t) c->part[t] = 0.0;
    c->u = u; c->v = v;
    par_for(c->N, dot_task, c);
    double s = 0.0;
    for (int t = 0; t < T; ++t) s += c->part[t];
    return s;
}

static void xr_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    pcg_ctx* c = (pcg_ctx*)c_;
    for (long long i = lo; i < hi; ++i) { c->x[i] += c->alpha * c->p[i]; c->r[i] -= c->alpha * c->Ap[i]; }
}
static void pz_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    pcg_ctx* c = (pcg_ctx*)c_;
    for (long long i = lo; i < hi; ++i) c->p[i] = c->z[i] + c->alpha * c->p[i];
}

/* per level, the vertices of each node in ascending order (counting sort) */
static void mas_node_lists(or_mas* m) {
    if (m->node_lists_ok) return;
    int n = m->n, L = m->L, tot = m->total;
    free(m->node_vptr); free(m->node_vlist);
    m->node_vptr = (int*)calloc((size_t)L * (tot + 1), sizeof(int));
    m->node_vlist = (int*)malloc(sizeof(int) * (size_t)L * (n > 0 ? n : 1));
    for (int l = 0; l < L; ++l) {
        int* ptr = m->node_vptr + (size_t)l * (tot + 1);
        int* lst = m->node_vlist + (size_t)l * n;
        for (int v = 0; v < n; ++v) ptr[or__node_of(m, v, l) + 1]++;
        for (int I = 0; I < tot; ++I) ptr[I + 1] += ptr[I];
        int* pos = (int*)malloc(sizeof(int) * (size_t)(tot + 1));
        memcpy(pos, ptr, sizeof(int) * (size_t)(tot + 1));
        for (int v = 0; v < n; ++v) lst[pos[or__node_of(m, v, l)]++] = v;
        free(pos);
    }
    m->node_lists_ok = 1;
}

static void restrict_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    pcg_ctx* c = (pcg_ctx*)c_;
    or_mas* m = c->mas;
    int tot = m->total, n = m->n;
    for (int l = 0; l < m->L; ++l) {
        const int* ptr = m->node_vptr + (size_t)l * (tot + 1);
        const int* lst = m->node_vlist + (size_t)l * n;
        for (long long I = lo; I < hi; ++I) {
            if (ptr[I] == ptr[I + 1]) continue;
            if (c->mas_fp32) {
                float s0 = 0, s1 = 0, s2 = 0;
                for (int e = ptr[I]; e < ptr[I + 1]; ++e) {
                    const double* rv = c->rin + 3 * (size_t)lst[e];
                    s0 += (float)rv[0]; s1 += (float)rv[1]; s2 += (float)rv[2];
                }
                c->Rf[3 * I] = s0; c->Rf[3 * I + 1] = s1; c->Rf[3 * I + 2] = s2;
            } else {
                double s0 = 0, s1 = 0, s2 = 0;
                for (int e = ptr[I]; e < ptr[I + 1]; ++e) {
                    const double* rv = c->rin + 3 * (size_t)lst[e];
                    s0 += rv[0]; s1 += rv[1]; s2 += rv[2];
                }
                c->R[3 * I] = s0; c->R[3 * I + 1] = s1; c->R[3 * I + 2] = s2;
            }
        }
    }
}
static void bank_solve_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    pcg_ctx* c = (pcg_ctx*)c_;
    or_mas* m = c->mas;
    int D = 3 * m->B;
    size_t bs = (size_t)D * D;
    for (long long b = lo; b < hi; ++b) {
        if (c->mas_fp32) {
            const float* Ai = m->AinvF + bs * b;
            for (int i = 0; i < D; ++i) {
                float s = 0;
                for (int j = 0; j < D; ++j) s += Ai[(size_t)i * D + j] * c->Rf[(size_t)b * D + j];
                c->Zf[(size_t)b * D + i] = s;
            }
        } else {
            const double* Ai = m->Ainv + bs * b;
            for (int i = 0; i < D; ++i) {
                double s = 0;
                for (int j = 0; j < D; ++j) s += Ai[(size_t)i * D + j] * c->R[(size_t)b * D + j];
                c->Z[(size_t)b * D + i] = s;
            }
        }
    }
}
static void collect_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    pcg_ctx* c = (pcg_ctx*)c_;
    or_mas* m = c->mas;
    for (long long v = lo; v < hi; ++v)
        for (int k = 0; k < 3; ++k) {
            if (c->mas_fp32) {
                float s = 0;
                for (int l = 0; l < m->L; ++l) s += c->Zf[3 * (size_t)or__node_of(m, (int)v, l) + k];
                c->zout[3 * v + k] = (double)s;
            } else {
                double s = 0;
                for (int l = 0; l < m->L; ++l) s += c->Z[3 * (size_t)or__node_of(m, (int)v, l) + k];
                c->zout[3 * v + k] = s;
            }
        }
}
static void diag_task(void* c_, int tid, long long lo, long long hi) {
    (void)tid;
    pcg_ctx* c = (pcg_ctx*)c_;
    for (long long i = lo; i < hi; ++i)
        for (int p = 0; p < 3; ++p) {
            double s = 0;
            for (int q = 0; q < 3; ++q) s += c->Dinv[9 * (size_t)i + p + 3 * q] * c->rin[3 * i + q];
            c->zout[3 * i + p] = s;
        }
}

static void apply_pre_mt(pcg_ctx* c, const double* r, double* z) {
    c->rin = r; c->zout = z;
    if (c->precond == 2 && c->mas && c->mas->n > 0) {
        or_mas* m = c->mas;
        mas_node_lists(m);
        /* slots of a bank that hold no node stay zero */
        if (c->mas_fp32) memset(c->Rf, 0, sizeof(float) * 3 * (size_t)m->total);
        else memset(c->R, 0, sizeof(double) * 3 * (size_t)m->total);
        par_for(m->total, restrict_task, c);
        par_for(m->total / m->B, bank_solve_task, c);
        par_for(m->n, collect_task, c);
        return;
    }
    if (c->precond == 1) { par_for(c->nb, diag_task, c); return; }
    memcpy(z, r, sizeof(double) * (size_t)c->N);
}

int or_pcg_mt(int K, const int* rows, const int* cols, const double* vals, int nb, const double* b, double* x,
              int precond, or_mas* mas, int mas_fp32, double tol, int rule, int max_iter, double* rel_res_out) {
    int T = pool.nthreads > 0 ? pool.nthreads : 1;
    pcg_ctx c;
    memset(&c, 0, sizeof(c));
    c.K = K; c.nb = nb; c.N = 3 * nb; c.rows = rows; c.cols = cols; c.vals = vals;
    c.precond = precond; c.mas = mas; c.mas_fp32 = mas_fp32;
    int N = c.N;
    /* row pointers and column lists of the upper BCOO (once per solve) */
    c.rptr = (int*)malloc(sizeof(int) * ((size_t)nb + 1));
    par_for(nb + 1, rowptr_task, &c);
    c.cnt = (long long*)calloc((size_t)T * (nb + 1), sizeof(long long));
    par_for(K, ccount_task, &c);
    c.cptr = (int*)malloc(sizeof(int) * ((size_t)nb + 1));
    long long acc = 0;
    for (int j = 0; j < nb; ++j) {
        c.cptr[j] = (int)acc;
        for (int t = 0; t < T; ++t) {
            long long v = c.cnt[(size_t)t * (nb + 1) + j];
            c.cnt[(size_t)t * (nb + 1) + j] = acc;
            acc += v;
        }
    }
    c.cptr[nb] = (int)acc;
    c.cidx = (int*)malloc(sizeof(int) * (size_t)(acc > 0 ? acc : 1));
    par_for(K, cfill_task, &c);
    c.part = (double*)calloc((size_t)T, sizeof(double));
    double *r = (double*)malloc(sizeof(double) * N), *z = (double*)malloc(sizeof(double) * N);
    double *p = (double*)malloc(sizeof(double) * N), *Ap = (double*)malloc(sizeof(double) * N);
    double* Dinv = NULL;
    if (precond == 1) {
        Dinv = (double*)calloc(9 * (size_t)nb, sizeof(double));
        for (int k = 0; k < K; ++k)
            if (rows[k] == cols[k]) {
                /* 3x3 inverse, column-major (the serial code's inv3) */
                const double* M = vals + 9 * (size_t)k;
                double* R = Dinv + 9 * (size_t)rows[k];
                double a = M[0], bq = M[3], cq = M[6], d = M[1], e = M[4], f = M[7], g = M[2], h = M[5], i = M[8];
                double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
                double id = 1.0 / (a * A + bq * B + cq * C);
                R[0] = A * id; R[3] = -(bq * i - cq * h) * id; R[6] = (bq * f - cq * e) * id;
                R[1] = B * id; R[4] = (a * i - cq * g) * id; R[7] = -(a * f - cq * d) * id;
                R[2] = C * id; R[5] = -(a * h - bq * g) * id; R[8] = (a * e - bq * d) * id;
            }
    }
    c.Dinv = Dinv;
    if (mas && mas->total > 0) {
        if (mas_fp32) {
            c.Rf = (float*)malloc(sizeof(float) * 3 * (size_t)mas->total);
            c.Zf = (float*)malloc(sizeof(float) * 3 * (size_t)mas->total);
        } else {
            c.R = (double*)malloc(sizeof(double) * 3 * (size_t)mas->total);
            c.Z = (double*)malloc(sizeof(double) * 3 * (size_t)mas->total);
        }
    }
    c.x = x; c.r = r; c.p = p; c.z = z; c.Ap = Ap;
    memset(x, 0, sizeof(double) * N);
    memcpy(r, b, sizeof(double) * N);
    double bn = sqrt(dot_mt(&c, b, b));
    apply_pre_mt(&c, r, z);
    double rz = dot_mt(&c, r, z), rz0 = rz;
    memcpy(p, z, sizeof(double) * N);
    int k;
    for (k = 1; k < max_iter; ++k) {
        spmv_mt(&c, 1.0, p, 0.0, Ap);
        double pAp = dot_mt(&c, p, Ap);
        c.alpha = rz / pAp;
        par_for(N, xr_task, &c);
        if (rule == 0) {
            if (fabs(rz) <= tol * rz0) break;
        } else {
            if (sqrt(dot_mt(&c, r, r)) <= tol * bn) break;
        }
        apply_pre_mt(&c, r, z);
        double rzn = dot_mt(&c, r, z);
        c.alpha = rzn / rz;  /* beta */
        par_for(N, pz_task, &c);
        rz = rzn;
    }
    if (rel_res_out) {
        spmv_mt(&c, 1.0, x, 0.0, Ap);
        double s = 0;
        for (int i = 0; i < N; ++i) { double d = b[i] - Ap[i]; s += d * d; }
        *rel_res_out = bn > 0 ? sqrt(s) / bn : sqrt(s);
    }
    free(r); free(z); free(p); free(Ap); free(Dinv);
    free(c.rptr); free(c.cnt); free(c.cptr); free(c.cidx); free(c.part);
    free(c.R); free(c.Z); free(c.Rf); free(c.Zf);
    return k;
}
