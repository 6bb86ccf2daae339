/*
 * oracle.h -- CPU restatement of the reference (jjacobx/qtnh) local contraction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (qtnh_amd/, include/) never links or
 * calls it.
 *
 * Elements are complex128 stored interleaved (re, im), the layout of
 * std::complex<double> (reference include/core/typedefs.hpp:21).  Tensors are
 * row-major over their dims (reference README.md:60, utils.cpp:27-35).
 *
 * Every function is pinned by a check in tests/test_oracle.py; see the
 * per-function comments for the reference lines each one restates and what
 * pins it.
 */
#ifndef QTNH_ORACLE_H
#define QTNH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_OK = 0, OR_EINVAL = 1, OR_ENOMEM = 2, OR_EUNSUPPORTED = 3 };

#define OR_X 0xFFFFu /* wildcard / "not placed" marker, reference typedefs.hpp:55 */

/* Row-major element count with a 64-bit accumulator (reference utils.cpp:23-25
 * accumulates in int; the restatement fixes that overflow on purpose). */
uint64_t or_size(int rank, const uint64_t *dims);

/* ---- seeded inputs: counter-based SplitMix64 + Box-Muller ---------------- */
/* SplitMix64 (Steele, Lea & Flood 2014): k-th output of the stream seeded by seed. */
uint64_t or_splitmix64(uint64_t seed, uint64_t k);
/* Natural log on (0, 1] from + - * / only (atanh series), deterministic across compilers. */
double or_log(double u);
/* sin and cos of 2*pi*v for v in [0, 1), octant reduction + Taylor polynomials. */
void or_sincos2pi(double v, double *s, double *c);
/* Fill n complex elements starting at element index `first`.
 * dist 0: re, im ~ N(0,1) by Box-Muller; dist 1: re, im uniform in [-1, 1). */
void or_fill_c128(double *out, uint64_t first, uint64_t n, uint64_t seed, int dist);

/* ---- data movement -------------------------------------------------------- */
/* Axis permutation (reference dense.cpp:152-169, 354-488): new_dims[ptup[i]] =
 * old_dims[i]; element at old multi-index (i_0..i_{r-1}) lands at the new
 * multi-index whose position ptup[k] holds i_k.  Out of place. */
int or_permute(const double *src, double *dst, int rank, const uint64_t *dims, const uint16_t *ptup);

/* In-place exchange of two equal-dimension axes (reference dense.cpp:171-218,
 * local case).  Returns OR_EINVAL for unequal dims (dense.cpp:178-180). */
int or_swap(double *data, int rank, const uint64_t *dims, int i, int j);

/* ---- contraction ----------------------------------------------------------- */
/* Default output placement (reference contract.cpp:219-248).  wires: nw pairs
 * (first in t1, second in t2).  repl1/repl2 receive, for every open axis, its
 * position in the result; wired axes get OR_X. */
int or_default_repls(int r1, int ndis1, int r2, int ndis2, int nw, const uint16_t *wires,
                     uint16_t *repl1, uint16_t *repl2);

/* Symmetric in-place placement (reference contract.cpp:250-289) for a t2 that
 * is a symmetric tensor with nd2 distributed and nl2 local axes and n_dis_in
 * distributed input axes.  Returns 1 and fills repl1/repl2 when the policy
 * applies, 0 when it does not (repl arrays untouched). */
int or_symm_repls(int r1, int nd2, int nl2, int n_dis_in, int nw, const uint16_t *wires,
                  uint16_t *repl1, uint16_t *repl2);

/* General pairwise contraction, the STEP-3 MAC loop of reference
 * contract.cpp:128-162 in global (one-rank) form: for each open t1 tuple
 * (row-major), each open t2 tuple (row-major), sum over the closed tuples in
 * the reference order (wires sorted by their t2 axis, the last one fastest;
 * indexing.cpp:27-35, 271-302), accumulator starting at 0, products as
 * libstdc++ std::complex<double> without FMA.  The result element goes to the
 * position given by repl1/repl2.  C must hold prod(open dims) elements. */
int or_contract(const double *A, int rA, const uint64_t *dA,
                const double *B, int rB, const uint64_t *dB,
                int nw, const uint16_t *wires,
                const uint16_t *repl1, const uint16_t *repl2, double *C);

/* Symmetric gate applied in place: the state's axes axes[m] are wired to the
 * gate's input axis m; the gate G has layout [in_0..in_{k-1}, out_0..out_{k-1}]
 * (reference symm.hpp:22-45).  Same arithmetic as or_contract under the
 * symmetric placement. */
int or_gate_apply(double *state, int rank, const uint64_t *dims, int k, const uint16_t *axes,
                  const double *G);

/* Diagonal gate in place: state[x] *= d[x restricted to axes] (the intended
 * DiagTensor semantics of reference diag.hpp:11-14, diag.cpp:19). */
int or_diag_apply(double *state, int rank, const uint64_t *dims, int k, const uint16_t *axes,
                  const double *d);

/* Quantum Fourier transform on an n-qubit state in the order of reference
 * tests/gen/qft.hpp:47-108: for each qubit i, H on i, then for j > i the
 * controlled phase exp(i*pi/2^(j-i)) on (i, j).  Qubit 0 is axis 0.  Phases
 * are taken from `phases[j-i]` as (re, im) pairs so the caller controls the
 * host math library. */
int or_qft(double *state, int n, const double *phases);

/* ---- CPU timing baseline (not checkers) ------------------------------------ */
/* The same per-element arithmetic as or_permute, or_gate_apply and
 * or_diag_apply, split over nthreads POSIX threads; results equal theirs bit
 * for bit. */
int or_permute_mt(const double *src, double *dst, int rank, const uint64_t *dims, const uint16_t *ptup, int nthreads);
int or_gate_apply_mt(double *state, int rank, const uint64_t *dims, int k, const uint16_t *axes, const double *G,
                     int nthreads);
int or_diag_apply_mt(double *state, int rank, const uint64_t *dims, int k, const uint16_t *axes, const double *d,
                     int nthreads);

/* ---- reductions used by property checks ------------------------------------ */
double or_norm2(const double *x, uint64_t n);
/* Position-independent multiset hash of the element bit patterns. */
uint64_t or_multiset_hash(const double *x, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
