/*
 * qtnh_hip.h -- C ABI of the MI355X (gfx950) kernels behind the qtnh local
 * contraction path.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types.  Element
 * type is complex128 stored interleaved (re, im), the layout of
 * std::complex<double> (reference include/core/typedefs.hpp:21).  Tensors are
 * row-major over their dims; counts are in elements (64-bit).
 *
 * Ownership: the caller owns every buffer; the library owns only the context
 * (device, stream, workspace).  Threading: a context is not thread-safe.  All
 * compute calls are stream-ordered and asynchronous on the context's stream;
 * qtnh_d2h and qtnh_sync block.  No call throws; every call returns a
 * qtnh_status and, on failure, leaves a message in qtnh_ctx_last_error().
 *
 * Each entry point names the reference interface it replaces (file:line in
 * jjacobx/qtnh).
 */
#ifndef QTNH_HIP_H
#define QTNH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  QTNH_OK = 0,
  QTNH_EINVAL = 1,       /* bad argument (shape, axis, wire, placement) */
  QTNH_ENOMEM = 2,       /* device allocation failed or budget exceeded */
  QTNH_EHIP = 3,         /* HIP runtime error (no device, launch failure) */
  QTNH_EUNSUPPORTED = 4  /* valid request outside what this build implements */
} qtnh_status;

/* Placement marker for wired axes (reference include/core/typedefs.hpp:55, X). */
#define QTNH_X 0xFFFFu
/* Largest tensor rank accepted (axes of any dims). */
#define QTNH_MAX_RANK 64

/* precision argument of qtnh_contract_c128 */
#define QTNH_PREC_FP64 0
#define QTNH_PREC_SPLIT 1

/* distribution argument of qtnh_fill_c128 */
#define QTNH_DIST_NORMAL 0
#define QTNH_DIST_UNIFORM 1

typedef struct qtnh_ctx_s *qtnh_ctx_t;

/* ---- context: replaces the MPI lifecycle of QTNHEnv (reference src/core/env.cpp:9-23) ---- */
/* stream: the hipStream_t every call of this context is ordered on (NULL: the
 * device's default stream).  The caller keeps ownership of the stream. */
qtnh_status qtnh_ctx_create(int device, void *stream, qtnh_ctx_t *out);
qtnh_status qtnh_ctx_destroy(qtnh_ctx_t ctx);
/* Releases the device and pinned buffers the context keeps between calls
 * (contraction workspace, descriptor, reduction and staging buffers) after its
 * stream has drained; later calls allocate them again as needed.  No reference
 * counterpart: for callers that hand large amounts of HBM to other work between
 * batches. */
qtnh_status qtnh_ctx_trim(qtnh_ctx_t ctx);
const char *qtnh_status_string(int status);
const char *qtnh_ctx_last_error(qtnh_ctx_t ctx);
/* Number of visible HIP devices (0 when none); never fails on a host without GPUs. */
qtnh_status qtnh_device_count(int *count);

/* ---- device storage: replaces std::vector<tel> loc_els_ (reference include/tensor/dense.hpp:24) ---- */
qtnh_status qtnh_buf_alloc(qtnh_ctx_t ctx, uint64_t n, void **out);
qtnh_status qtnh_buf_free(qtnh_ctx_t ctx, void *buf);
qtnh_status qtnh_h2d(qtnh_ctx_t ctx, void *dst, const void *src, uint64_t n);
qtnh_status qtnh_d2h(qtnh_ctx_t ctx, void *dst, const void *src, uint64_t n);
qtnh_status qtnh_d2d(qtnh_ctx_t ctx, void *dst, const void *src, uint64_t n);
qtnh_status qtnh_sync(qtnh_ctx_t ctx);

/* Seeded on-device fill (no reference counterpart: synthetic inputs of the
 * benchmarks).  Element g uses SplitMix64 outputs 2g and 2g+1 of `seed`;
 * dist QTNH_DIST_NORMAL gives re, im ~ N(0,1) by Box-Muller. */
qtnh_status qtnh_fill_c128(qtnh_ctx_t ctx, void *buf, uint64_t n, uint64_t seed, int dist);

/* ---- output placement: host logic of Tensor::contract (reference src/tensor/contract.cpp:219-289) ---- */
/* Default placement (contract.cpp:219-248).  wires: nw (t1 axis, t2 axis) pairs. */
qtnh_status qtnh_default_repls(int r1, int ndis1, int r2, int ndis2, int nw, const uint16_t *wires,
                               uint16_t *repl1, uint16_t *repl2);
/* Symmetric in-place placement (contract.cpp:250-289) for a symmetric t2 with
 * nd2 distributed axes, nl2 local axes and n_dis_in distributed input axes.
 * *applies is set to 1 (repls written) or 0 (policy does not apply). */
qtnh_status qtnh_symm_repls(int r1, int nd2, int nl2, int n_dis_in, int nw, const uint16_t *wires,
                            uint16_t *repl1, uint16_t *repl2, int *applies);
/* Result dims of a contraction under the given placement; validates that the
 * placement is a permutation of the result axes (contract.cpp:86-111). */
qtnh_status qtnh_contract_dims(int rA, const uint64_t *dA, int rB, const uint64_t *dB, int nw,
                               const uint16_t *wires, const uint16_t *repl1, const uint16_t *repl2,
                               int *rC, uint64_t *dC);

/* ---- data movement ---- */
/* Axis permutation, out of place: dst axis ptup[i] = src axis i.
 * Replaces DenseTensor::permute / TIDense::_permute_internal (reference
 * src/tensor/dense.cpp:152-169, 354-488; Tensor::permute tensor.hpp:190-192). */
qtnh_status qtnh_permute_c128(qtnh_ctx_t ctx, const void *src, void *dst, int rank, const uint64_t *dims,
                              const uint16_t *ptup);
/* In-place exchange of two equal-dim axes.  Replaces DenseTensor::swap /
 * TIDense::_swap_internal local case (reference src/tensor/dense.cpp:108-111, 171-218). */
qtnh_status qtnh_swap_c128(qtnh_ctx_t ctx, void *buf, int rank, const uint64_t *dims, int i, int j);

/* ---- gate application (the symmetric in-place policy of contract.cpp:250-289 + STEP 3 :131-162) ---- */
/* In place: psi'[.. o@axes ..] = sum_i psi[.. i@axes ..] * G[i, o], G host
 * array of prod(d)^2 complex, layout [in_0..in_{k-1}, out_0..out_{k-1}]
 * (reference include/tensor/symm.hpp:22-45).  Sum over i in row-major order
 * from 0, products without FMA: the reference's arithmetic, bit for bit.
 * Replaces Tensor::contract(state, SymmTensor) with all gate inputs wired.
 * flags: reserved, pass 0. */
qtnh_status qtnh_gate_apply_c128(qtnh_ctx_t ctx, void *state, int rank, const uint64_t *dims, int k,
                                 const uint16_t *axes, const double *G, int flags);
/* In place: psi[x] *= d[x@axes], d host array of prod(d) complex.  The
 * intended DiagTensor contraction (reference include/tensor/diag.hpp:11-14,
 * 63-65; src/tensor/diag.cpp:19). */
qtnh_status qtnh_diag_apply_c128(qtnh_ctx_t ctx, void *state, int rank, const uint64_t *dims, int k,
                                 const uint16_t *axes, const double *d);

/* ---- general pairwise contraction: _contract_dense (reference src/tensor/contract.cpp:12-201) ---- */
/* C = sum over wires of A*B, result placed by repl1/repl2 (QTNH_X for wired
 * axes).  Small problems use a kernel that keeps the reference summation
 * order; large ones permute to matrices and run a complex GEMM on MFMA.
 * precision QTNH_PREC_FP64 (the drop-in) or QTNH_PREC_SPLIT: GEMM-sized
 * problems then run as exact INT8 matrix-core products of 8-bit slices of A
 * and B recombined in FP64 (finite inputs only).  err_out (may be NULL)
 * receives eta of the componentwise a-priori bound
 *   |C_ij - C~_ij| <= eta * max_k|A_ik| * max_k|B_kj|   (max over re and im)
 * for SPLIT, and 0 when the FP64 kernels ran.  A, B and C must not overlap. */
qtnh_status qtnh_contract_c128(qtnh_ctx_t ctx, const void *A, int rA, const uint64_t *dA, const void *B,
                               int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                               const uint16_t *repl1, const uint16_t *repl2, void *C, int precision,
                               double *err_out);

/* Bytes of the static device scratch of qtnh_contract_c128_noalloc (per GPU). */
#define QTNH_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* qtnh_contract_c128 without any device allocation.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Same
 * arguments, routes, kernels and results as qtnh_contract_c128; the GEMM
 * route's workspace (permuted operands, natural result, split-precision digit
 * planes, split-K and depth-chunk partials) is a file-scope device array of
 * QTNH_NOALLOC_SCRATCH_BYTES that every GPU holds until the process ends, its
 * address looked up once per device by qtnh_ctx_create.  A contraction whose
 * workspace exceeds it returns QTNH_EINVAL before any launch.  Kernels that
 * time their tile shapes on first use keep the cached or the default shape
 * during the call (no events, no synchronisation; bits do not depend on the
 * shape).  Calls that overlap in time on one GPU (other streams, other
 * threads) share the array and race; calls ordered on one stream do not. */
qtnh_status qtnh_contract_c128_noalloc(qtnh_ctx_t ctx, const void *A, int rA, const uint64_t *dA, const void *B,
                                       int rB, const uint64_t *dB, int nw, const uint16_t *wires,
                                       const uint16_t *repl1, const uint16_t *repl2, void *C, int precision,
                                       double *err_out);

/* ---- reductions used for normalisation and checks ---- */
/* *out = sum |x_i|^2 (device reduction, fixed order for a given n). */
qtnh_status qtnh_norm2_c128(qtnh_ctx_t ctx, const void *x, uint64_t n, double *out);
/* x_i *= (re + i im) */
qtnh_status qtnh_scale_c128(qtnh_ctx_t ctx, void *x, uint64_t n, double re, double im);

/* ---- circuits: gate sequences on a resident state (no single reference
 * function: the reference applies a circuit as one Tensor::contract per gate,
 * examples/qft.cpp:35-111, tests/gen/qft.hpp:10-111, through the symmetric
 * in-place placement of contract.cpp:250-289).  Qubit q is axis q of a rank-n
 * state with every dim 2 (axis 0 most significant). ---- */
typedef struct qtnh_circuit_s *qtnh_circuit_t;
#define QTNH_RUN_PER_GATE 0 /* one streaming launch per op (the reference's one contraction per gate) */
#define QTNH_RUN_FUSED 1    /* ops grouped into runs of <= tile-bits qubits, one state pass per run */

qtnh_status qtnh_circuit_create(int n_qubits, qtnh_circuit_t *out);
qtnh_status qtnh_circuit_destroy(qtnh_circuit_t c);
/* Appends a dense k-qubit gate (k <= 3), G[in][out] row-major, (2^k)^2
 * complex, contracted over `in` like qtnh_gate_apply_c128. */
qtnh_status qtnh_circuit_gate(qtnh_circuit_t c, int k, const uint16_t *qubits, const double *G);
/* Appends a diagonal k-qubit gate (k <= 3), d of 2^k complex, like qtnh_diag_apply_c128. */
qtnh_status qtnh_circuit_diag(qtnh_circuit_t c, int k, const uint16_t *qubits, const double *d);
/* Appends the QFT of the reference's example and tests (tests/gen/qft.hpp:47-108):
 * for i: H on i, then for j > i the controlled phase diag(1,1,1,e^{i pi/2^(j-i)})
 * on (i, j).  phases[2m], phases[2m+1] hold e^{i pi/2^m} (m = 1..n-1); NULL
 * computes them with cos/sin. */
qtnh_status qtnh_circuit_qft(qtnh_circuit_t c, const double *phases);
/* Appends `depth` layers of a seeded random circuit: a Haar-random SU(2) on
 * every qubit, then CZ on the pairs of a seeded random pairing of the qubits. */
qtnh_status qtnh_circuit_random(qtnh_circuit_t c, int depth, uint64_t seed);
/* Appends `depth` layers on a rows x cols grid (qubit r*cols + c; rows*cols
 * must equal the circuit's qubits): a Haar-random SU(2) on every qubit, then
 * CZ on grid neighbours, cycling horizontal-even, horizontal-odd,
 * vertical-even, vertical-odd couplers layer by layer. */
qtnh_status qtnh_circuit_grid(qtnh_circuit_t c, int rows, int cols, int depth, uint64_t seed);
/* Number of qubits, ops, and runs under the circuit's tuned tile (qtnh_circuit_tune), else the
 * process default tile (any may be NULL). */
qtnh_status qtnh_circuit_info(qtnh_circuit_t c, int *n_qubits, uint64_t *n_ops, uint64_t *n_runs);
/* Reads op i back: kind (0 gate, 1 diagonal), k, qubits[k], coefficients. */
qtnh_status qtnh_circuit_op(qtnh_circuit_t c, uint64_t i, int *kind, int *k, uint16_t *qubits, double *coef);
/* Applies the circuit in place to a 2^n-amplitude state; both modes give
 * bit-identical results.  QTNH_RUN_FUSED keeps the run descriptors it uploads
 * for each state buffer (up to 64 per circuit): a later fused run of the same
 * circuit on the same buffer only launches kernels, so it may be captured
 * into a HIP graph; the graph stays valid while the circuit is not modified
 * or destroyed (tuning calls and runs on other devices recompile the circuit
 * but keep the device memory earlier graphs read until then).  A fused run that would upload (first run on a buffer, or a
 * circuit not yet compiled) returns QTNH_EUNSUPPORTED while the context's
 * stream is being captured. */
qtnh_status qtnh_circuit_run(qtnh_ctx_t ctx, qtnh_circuit_t c, void *state, int mode);
/* Applies circuit i to states[i] (distinct buffers) for every i, fused, with
 * one launch per run index across the whole batch. */
qtnh_status qtnh_circuit_run_batch(qtnh_ctx_t ctx, int count, const qtnh_circuit_t *circuits, void *const *states);

/* ---- tensor networks: the caller of the contraction path
 * (TensorNetwork::contractBond / contractAll, reference src/tensor/network.cpp:60-170).
 * Leaf axes carry int64 bond labels; a label shared by two axes is a wire.
 * Contractions use the default placement (open axes of the first tensor,
 * then of the second; contract.cpp:219-248). ---- */
typedef struct qtnh_net_s *qtnh_net_t;
qtnh_status qtnh_net_create(qtnh_net_t *out);
qtnh_status qtnh_net_destroy(qtnh_net_t net);
/* Adds a leaf tensor (host data, row-major complex); *id (may be NULL) gets its node id. */
qtnh_status qtnh_net_add_tensor(qtnh_net_t net, int rank, const uint64_t *dims, const int64_t *bonds,
                                const double *data, int *id);
/* The best of `tries` greedy plans by estimated time.  A plan's cost is the
 * sum over its steps of 8 M N K flops + 1024 x the elements the step reads
 * and writes (operands and result): flops + 64 x the bytes qtnh_net_info
 * reports.  Try 0 is qtnh_net_plan's order for `seed`, try t >= 1 a sampled
 * greedy order (random-greedy at low temperature) drawn from seed + t; each
 * try contracts its last 10 tensors in their cheapest order (exhaustive
 * search over subsets).  Keeps the least cost among the tries whose
 * intermediates stay within max_rank (ties: the lowest try), and plans the
 * kept try with the search over its last 12 tensors (never a higher cost);
 * *chosen_seed (may be NULL) gets seed + t of the kept try.  Networks of at
 * most 12 tensors get their cheapest order.  The tries run on up to 16 host
 * threads; the result does not depend on their number.  QTNH_EUNSUPPORTED
 * when no try fits the cap. */
qtnh_status qtnh_net_plan_best(qtnh_net_t net, uint64_t seed, int max_rank, int tries, uint64_t *chosen_seed);
/* Adds a leaf whose elements stay in caller device memory (row-major complex,
 * read in place by every run; it must stay valid and unchanged until the runs
 * that read it have finished).  bonds may be NULL when the network is planned
 * with qtnh_net_plan_steps, which does not use labels. */
qtnh_status qtnh_net_add_tensor_device(qtnh_net_t net, int rank, const uint64_t *dims, const int64_t *bonds,
                                       const void *data, int *id);
/* Builds the amplitude network <bits| C |0...0> of a circuit: |0> vectors,
 * one tensor per op (axes: inputs, then outputs; diagonals as dense tensors),
 * and <bits_q| vectors (bits NULL: all zeros).  The network must be empty. */
qtnh_status qtnh_net_from_circuit(qtnh_net_t net, qtnh_circuit_t c, const uint8_t *bits);
/* Greedy pairwise order (seeded ties) with fixed arena offsets for every
 * intermediate.  QTNH_EUNSUPPORTED when an intermediate would exceed max_rank. */
qtnh_status qtnh_net_plan(qtnh_net_t net, uint64_t seed, int max_rank);
/* One step of a caller-given order: contract node a with node b over nw wires
 * (axis of a, axis of b) and place the result by repl1 / repl2 (the result
 * position of each axis of a and of b, QTNH_X for wired axes; both NULL: the
 * default placement, open axes of a then of b). */
typedef struct {
  int a, b;
  int nw;
  const uint16_t *wires;
  const uint16_t *repl1;
  const uint16_t *repl2;
} qtnh_net_step_spec;
/* Plans the network in a caller-given order instead of the greedy one: leaves
 * are nodes 0..L-1 and the result of step s is node L+s; there are L-1 steps,
 * each consuming two live nodes, so the last step's result is the network's.
 * Intermediates get fixed arena offsets as with qtnh_net_plan.  The order of
 * TensorNetwork::contractAll(order) / contractBond (reference
 * src/tensor/network.cpp:60-170), whose steps place their results as
 * Tensor::contract does (contract.cpp:219-289). */
qtnh_status qtnh_net_plan_steps(qtnh_net_t net, int nsteps, const qtnh_net_step_spec *steps);
/* Arithmetic of the network's large (GEMM-route) steps: QTNH_PREC_FP64 (default)
 * or QTNH_PREC_SPLIT (the split-precision GEMM of qtnh_contract_c128, for steps
 * whose result fills the GPU with at least 256 tiles of 128 x 128 over the
 * M x 2N real embedding; narrower steps stay FP64); the small steps always run
 * exactly in FP64 in the reference order.  Before or after
 * planning.  A prepared batch (qtnh_net_batch_create) keeps the precisions its
 * networks had when it was created.  No reference counterpart (the reference
 * contracts in complex<double> only). */
qtnh_status qtnh_net_set_precision(qtnh_net_t net, int precision);
/* Plan summary: leaves, steps, largest rank, sum of 8 M N K, sum of
 * 16 (|A| + |B| + |C|) bytes, device bytes of one run (any may be NULL). */
qtnh_status qtnh_net_info(qtnh_net_t net, uint64_t *n_tensors, uint64_t *n_steps, int *max_rank, double *flops,
                          double *bytes, uint64_t *device_bytes);
qtnh_status qtnh_net_result_dims(qtnh_net_t net, int *rank, uint64_t *dims);
/* Step s of the plan: node ids of its operands and result, and whether the
 * small-contraction kernel (reference summation order) computes it. */
qtnh_status qtnh_net_step(qtnh_net_t net, uint64_t s, int *a, int *b, int *c, int *direct);
/* Contracts the planned network, one launch per step; out (host) receives the
 * result (out_cap complex elements at least).  Blocks. */
qtnh_status qtnh_net_run(qtnh_ctx_t ctx, qtnh_net_t net, double *out, uint64_t out_cap);
/* Contracts the planned network on the context's stream into device buffer
 * out (the result's element count), without blocking: intermediates live in
 * the context's workspace, the last step writes out directly.  out must not
 * alias a leaf. */
qtnh_status qtnh_net_run_device(qtnh_ctx_t ctx, qtnh_net_t net, void *out);
/* Contracts planned networks together: step k of every network in grouped
 * launches (small binary contractions; small FP64 GEMM-route steps as grouped
 * operand permutes, GEMM tiles and split-K sums) plus the large steps in turn.  Networks
 * are admitted in consecutive waves whose device blocks fit budget_bytes.
 * outs[i] (host) receives network i's result; results are bit-identical to
 * qtnh_net_run.  *waves (may be NULL) gets the number of waves.  Blocks. */
qtnh_status qtnh_net_run_batch(qtnh_ctx_t ctx, int count, const qtnh_net_t *nets, double *const *outs,
                               uint64_t budget_bytes, int *waves);
/* A prepared batch for repeated evaluation (same structures, new leaf values):
 * device memory, descriptors and the launch sequence are built once; with
 * use_graph the sequence is captured into a HIP graph (on a private stream)
 * and each run is one graph launch on the context's stream.  The networks
 * must outlive the batch and must not be re-planned. */
typedef struct qtnh_net_batch_s *qtnh_net_batch_t;
qtnh_status qtnh_net_batch_create(qtnh_ctx_t ctx, int count, const qtnh_net_t *nets, int use_graph,
                                  qtnh_net_batch_t *out);
/* New values (host, same size) for leaf `leaf` of network `net`; stream-ordered. */
qtnh_status qtnh_net_batch_set_leaf(qtnh_ctx_t ctx, qtnh_net_batch_t b, int net, int leaf, const double *data);
/* Contracts every network of the batch; outs[i] (host) receives network i's result.  Blocks. */
qtnh_status qtnh_net_batch_run(qtnh_ctx_t ctx, qtnh_net_batch_t b, double *const *outs);
/* How the batch's steps run (any may be NULL): small direct steps in one
 * grouped launch per step index; small FP64 GEMM-route steps in the step
 * index's grouped permute / GEMM / split-K launches; the rest one launch
 * sequence each. */
qtnh_status qtnh_net_batch_stats(qtnh_net_batch_t b, uint64_t *grouped_direct, uint64_t *grouped_gemm,
                                 uint64_t *single);
qtnh_status qtnh_net_batch_destroy(qtnh_net_batch_t b);

/* ---- serving loop (no reference counterpart): queries on the output states of
 * seeded random SU(2) + CZ circuits (qtnh_circuit_random), answered on the
 * wall clock under an HBM budget ---- */
#define QTNH_QUERY_AMPLITUDE 0 /* value = <basis| C |0...0> */
#define QTNH_QUERY_EXPECT_Z 1  /* value = (<Z_0>, <Z_{n-1}>), zsum = sum_q <Z_q> */
#define QTNH_SERVE_SEQUENTIAL 0 /* one query at a time, one launch per gate, no resident states */
#define QTNH_SERVE_BATCHED 1    /* continuous batching: steps of grouped fused runs over every circuit
                                   in flight; arrivals admitted FIFO at each step while their states
                                   fit the budget; least-recently-used resident output states */
#define QTNH_SERVE_SEQUENTIAL_FUSED 2 /* one query at a time, fused runs, no resident states */
#define QTNH_SERVE_NETWORK 3    /* as BATCHED, but each amplitude query takes the cheaper route by
                                   estimated time: its state vector, or its tensor network <x|C|0>
                                   (best of four greedy orders within rank 28), contracted
                                   with the step's other network queries in one grouped batch */
/* Flags for QTNH_SERVE_NETWORK (or-ed into the policy): */
#define QTNH_SERVE_SPLIT 0x100       /* network GEMM steps that fill the GPU in split precision */
#define QTNH_SERVE_ALL_NETWORK 0x200 /* every amplitude query whose plan fits takes the network route */
typedef struct {
  int n_qubits;
  int depth;
  uint64_t seed;
  int kind;
  uint64_t basis;   /* amplitude index (qubit 0 most significant) */
  double arrival_s; /* seconds after the start of qtnh_serve */
} qtnh_query;
typedef struct {
  double start_s, done_s; /* seconds after the start */
  double value[2];
  double zsum;
  int batch;  /* batched: step at which it was admitted; sequential: its position in service order */
  int cached; /* 1: answered from a resident state */
  int route;  /* 0: state vector, 1: tensor network */
} qtnh_query_result;
/* Serves the queries in arrival order; returns when all are answered.  A
 * query's answer does not depend on the queries served with it; state-vector
 * answers are identical under every policy. */
qtnh_status qtnh_serve(qtnh_ctx_t ctx, int nq, const qtnh_query *queries, qtnh_query_result *results,
                       uint64_t budget_bytes, int policy);

/* Times the fused executor at several tile shapes on a depth-6 random
 * circuit of min(n, 26) qubits and keeps the fastest for n-qubit circuits
 * run on this context (results do not depend on it).  Outputs may be NULL:
 * the chosen tile bits, low bits and the best time (ms). */
qtnh_status qtnh_autotune_circuit(qtnh_ctx_t ctx, int n_qubits, int *tile_bits, int *low_bits, double *ms);
/* Times this circuit's own fused runs at several tile shapes (12 bits down to
 * 7) on `scratch`, a device buffer of 2^n complex128 elements whose contents
 * it overwrites, and keeps the fastest for every later fused run of the
 * circuit (any state, any context; results do not depend on it).  Worth it
 * for circuits run many times: the best tile depends on the circuit (QFT-12:
 * 8 bits, 65 against 77 us per run), not only on its size.  Outputs may be
 * NULL: tile bits, low bits, time per run (us).  No reference counterpart. */
qtnh_status qtnh_circuit_tune(qtnh_ctx_t ctx, qtnh_circuit_t c, void *scratch, int *tile_bits, int *low_bits,
                              double *us);

/* ---- queries on a state (no reference counterpart) ---- */
/* out[j] = <psi| Z_{qubits[j]} |psi> = sum_x |psi_x|^2 (1 - 2 x_q), fixed
 * reduction order; blocks until the values are on the host. */
qtnh_status qtnh_expect_z_c128(qtnh_ctx_t ctx, const void *state, int n_qubits, int nq, const uint16_t *qubits,
                               double *out);

/* ---- tuning (no reference counterpart) ---- */
/* Process-wide tile parameters of the power-of-two permute kernel: starting
 * source-low / result-low tile bits a, b (1..12), target tile bits (<= 12),
 * tile walk order (0: source order, 1: page-local on both sides; + 2: a tile
 * short of the target is filled with outer bits instead of longer runs; + 4:
 * tiles walked by their outer bits' distance from bit 10),
 * non-temporal loads and stores (0/1), persistent blocks per compute unit.  Results do not
 * depend on them.  Setting them turns per-permutation tuning off. */
void qtnh_tune_permute(int a, int b, int u_target, int order, int nt, int blocks_per_cu);
/* Per-permutation tuning of the power-of-two permute (on by default): the
 * first permute of >= 2^23 elements with a new bit permutation on a device
 * times a few tile shapes on the caller's buffers (one stream synchronisation;
 * skipped while the stream is being captured) and keeps the fastest.  on = 1
 * restores the default tile and clears the tuned shapes; 0 uses the set tile
 * for every permutation.  Results do not depend on it. */
void qtnh_tune_permute_auto(int on);
/* Block width (64 or 128 complex columns, 0: chosen by size) of the FP64
 * complex GEMM. */
void qtnh_tune_zgemm(int bn);
/* Number of 8-bit slices (1..8, default 5) of QTNH_PREC_SPLIT contractions. */
void qtnh_tune_split(int slices);
/* Block tile of the split-precision GEMM (up to 5 slices): 0 (default)
 * 128 x 128, one workgroup per compute unit; 1 64 x 64, two per compute unit;
 * 2 128 x 64; 3 64 x 64 with eight narrow waves.  Results do not depend on it. */
void qtnh_tune_split_tile(int shape);
/* Most workgroups (256..2^22, default 2^20) of the streaming gate kernel for
 * gates on high bits, which loops over the rest.  (The diagonal kernel and the
 * kernels for gates on the lowest 6 bits size their grids to the state.) */
void qtnh_tune_stream(int max_blocks);
/* Tile bits (3..13, default 12) of fused circuit runs, and how many of the
 * lowest state bits every tile keeps (0..8, default 4; 2^low x 16 bytes are
 * contiguous in each tile row).  Results do not depend on them. */
void qtnh_tune_circuit(int tile_bits, int low_bits);
/* Kernel for large states with 12-bit tiles: 2 (default) register-blocked
 * (ops applied to 16 amplitudes per thread in registers, one LDS round trip
 * per phase of ops on 4 tile bits), 3 register-blocked with the next tile
 * streamed into a second LDS buffer during the current tile's phases, 1
 * pipelined (per-op LDS passes), 0 plain.  Results do not depend on it. */
void qtnh_tune_circuit_pipeline(int mode);

#ifdef __cplusplus
}
#endif
#endif
