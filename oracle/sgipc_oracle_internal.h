/*
 * sgipc_oracle_internal.h -- state shared by the oracle's serial code
 * (sgipc_oracle.c) and its multi-threaded CPU-baseline variants
 * (sgipc_oracle_mt.c).  TEST INFRASTRUCTURE ONLY (see sgipc_oracle.h).
 */
#ifndef SGIPC_ORACLE_INTERNAL_H
#define SGIPC_ORACLE_INTERNAL_H
#include "sgipc_oracle.h"

struct or_mas {
    int n, M, B, L, total, nbanks_alloc;
    int *part_map, *real_map;
    uint32_t *nstart, *nnum0, *nlist0, *nnum, *nlist;
    int nlist_len;
    int* level_size; /* 2*(L+1) : (count, begin) */
    int* coarse_space; /* L*n */
    int* going_next; /* cap */
    int going_cap;
    int* coarse_table; /* 6*n */
    uint32_t* fine_mask; /* n */
    int* prefix0; /* M/B */
    double* A; /* banks * (3B)^2 dense Galerkin matrices */
    double* Ainv; /* symmetrised inverse */
    float* AinvF;
    int fp32_store;
    /* per level: the vertices of each node, ascending (multi-threaded apply;
     * built on first use after a build, freed with the MAS) */
    int* node_vptr; /* L * (total + 1) */
    int* node_vlist; /* L * n */
    int node_lists_ok;
};

/* threads of the multi-threaded variants (1: the serial code runs) */
extern int or_threads;

/* serial pieces the threaded code shares */
int or__node_of(const or_mas* m, int v, int l);
void or__invert_bank(or_mas* m, int b, double* S, double* colm);

/* threaded variants (or_threads > 1), same contracts as the serial ones */
void or_fem_assemble_mt(int n, int T, const double* x, const uint32_t* tets, const double* DmInv, const double* vol,
                        const double* mu, const double* lam, double dt, const int* btype, int off, double* grad,
                        double* vals, int* rows, int* cols);
int or_convert_mt(int nt, const int* rows, const int* cols, const double* vals, int* orow, int* ocol, double* oval);
void or_mas_project_invert_mt(or_mas* m, int K, const int* rows, const int* cols, const double* vals, int offset);
int or_pcg_mt(int K, const int* rows, const int* cols, const double* vals, int nb, const double* b, double* x,
              int precond, or_mas* mas, int mas_fp32, double tol, int rule, int max_iter, double* rel_res_out);
#endif
