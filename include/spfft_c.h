/* spfft_c.h -- C ABI of the MI355X (gfx950) binary-sparse 2-D R2C FFT.
 *
 * Plain C types only (pointers, ints, sizes); device pointers are HIP device
 * addresses and `stream` is a hipStream_t passed as void* (NULL = the legacy
 * default stream).  Every function returns an spfft_status; on failure
 * spfft_last_error() returns a message for the calling thread.
 *
 * Output layout (all entry points): X[u][v] = sum over nonzeros (r, c) of
 * exp(-2 pi i (u r / rows + v c / cols)), u < rows, v <= cols/2, complex64
 * row-major with pitch P = 4 * ceil((cols/2 + 1) / 4) elements -- the layout
 * of the reference's SparseFFTResult (src/sparse_fft_bluestein.cu:2259-2260,
 * :2391).  Padding columns v in [cols/2+1, P) are not written.
 */
#ifndef SPFFT_C_H
#define SPFFT_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum spfft_status {
    SPFFT_OK = 0,
    SPFFT_ERR_INVALID_ARG = 1, /* bad size, tile, option or NULL pointer       */
    SPFFT_ERR_HIP = 2,         /* a HIP runtime call failed                    */
    SPFFT_ERR_ALLOC = 3,       /* device or host allocation failed             */
    SPFFT_ERR_UNSUPPORTED = 4, /* shape outside what the kernels support       */
    SPFFT_ERR_BAD_INPUT = 5    /* COO index outside [0, rows) x [0, cols)      */
} spfft_status;

/* Bluestein length rule.  SMOOTH_3: M = next 2^a 3^b >= 2*cols-1 (the
 * reference's --csc-mixed-radix, src/sparse_fft_bluestein.cu:2261).
 * SMOOTH_7: M = next 2^a 3^b 5^c 7^d >= 2*cols-1 (--csc-cufft-smooth, :2562). */
enum { SPFFT_SMOOTH_3 = 3, SPFFT_SMOOTH_7 = 7 };

/* Arithmetic form of the length-m1 sub-FFTs of the row pass.
 * FP32: register codelets in FP32 (the plain form).
 * BF16X3: the radix-16/32 butterflies as complex DFT-matrix products on the
 * matrix cores (v_mfma_f32_32x32x16_bf16), BF16 hi/lo split of both operands,
 * three products accumulated in FP32; m1 is 16 or 32, and with fft_len = 0 the
 * Bluestein length is the smallest smooth M >= 2*cols-1 that splits as m1 * m2
 * with a register-codelet convolution kernel for m2. */
enum { SPFFT_PREC_FP32 = 0, SPFFT_PREC_BF16X3 = 1 };

/* The structs that grow with the library -- spfft_opts, spfft_plan_info and
 * spfft_result -- start with struct_size = sizeof(struct) as the caller's header
 * declares it: spfft_opts_default sets it for spfft_opts, and a caller sets it in
 * spfft_plan_info / spfft_result before asking for one.  A library built against
 * another layout of the struct rejects the call with SPFFT_ERR_INVALID_ARG instead
 * of reading or writing past the end of the caller's struct. */
typedef struct spfft_opts {
    int struct_size; /* sizeof(spfft_opts)                                    */
    int length_rule; /* SPFFT_SMOOTH_3 or SPFFT_SMOOTH_7                      */
    int u_tile;      /* output rows per tile (reference --csc-tile); 0: one tile of all
                        rows/2 + 1 computed rows                                   */
    int precision;   /* SPFFT_PREC_*                                           */
    int fft_len;     /* 0: from length_rule; else an explicit 7-smooth M >= 2*cols-1 */
    int m1;          /* 0: automatic; else the first factor of M = m1 * m2 (a 7-smooth m1 <= 64);
                        row_transform SPLIT: the split's n1 (cols = n1 * n2)               */
    int conv_rows;   /* 0: automatic; else rows per workgroup of the convolution kernel (<= 64) */
    int colpass_rows;  /* 0: automatic; else 1, 2, 4, 8 or 16 output rows per lane in the column pass */
    int colpass_table; /* 0: automatic; 1: two-level W_R table; 2: whole W_R table in LDS      */
    int row_limit;     /* 0: all rows.  Else only output rows u < row_limit (and their mirror
                          rows) are computed: a sampled run for timing (the autotuner) */
    int row_kernel;    /* SPFFT_ROW_*: how the Bluestein row pass runs                       */
    int colpass_pair;  /* SPFFT_PAIR_*: half-period row pairing in the column pass            */
    int row_transform; /* SPFFT_TRANSFORM_*: Bluestein, a direct length-cols FFT, or Rader     */
    int row_persist;   /* fused row kernel: 0 persistent (the resident workgroups loop over the
                          rows and load the twiddle tables once), 1 one workgroup per row group */
    int t_rows;        /* four-step kernels: rows of the tile buffer T (t_rows x M complex64);
                          the column pass fills a u_tile of rows once, then k_fft1 / k_conv /
                          k_ifft1 run over it t_rows at a time.  0: the whole tile.  (Sub-tiles
                          small enough for T to stay in the 256 MB last-level cache measured
                          slower on the 40421^2 matrix: 96-512 rows 18.6-21.5 ms against
                          18.0 ms for one tile.) */
    int row_threads;   /* on-chip row kernels (fused, direct, Rader): 0 the first listed shape for
                          the length; else only a shape of this many threads per workgroup */
} spfft_opts;

/* Row-pass kernels.  FOUR_STEP: k_fft1 + k_conv + k_ifft1 with the tile buffer T
 * in HBM (any supported M).  FUSED: one kernel per row with the length-M row
 * held in registers and LDS (no T; the M of a fixed list of shapes, FP32 form).
 * AUTO: FUSED when M has a fused shape and the options allow it, else FOUR_STEP. */
enum { SPFFT_ROW_AUTO = 0, SPFFT_ROW_FOUR_STEP = 1, SPFFT_ROW_FUSED = 2 };

/* Column-pass row pairing.  For even rows R, W_R^{(u + R/2) r} = (-1)^r W_R^{u r}:
 * one phase lookup per nonzero serves row u (even-r sum + odd-r sum) and row
 * R/2 - u (the conjugate of their difference), so the pass computes rows
 * u <= R/4 only.  AUTO: paired when R is even, one tile holds every computed row
 * (u_tile >= R/2 + 1), row_limit is 0 and colpass_rows is not 1 (a lane's output
 * rows are then half its own, half their partners).  OFF: every row on its own. */
enum { SPFFT_PAIR_AUTO = 0, SPFFT_PAIR_OFF = 1 };

/* Row transform.  BLUESTEIN: the chirp convolution at the smooth length M >= 2 cols - 1
 * of the reference's variants.  DIRECT: one FFT of length cols itself when cols has a
 * fused-kernel shape holding the whole row on chip (cols in spfft_direct_lengths();
 * FP32 form, no fft_len / conv_rows / four-step options): no chirps, a quarter of the
 * row-pass work.  RADER: for cols = 2p or cols = p, p an odd prime with a Rader shape
 * for p - 1 (cols in spfft_rader_widths(); same restrictions): the length-p DFTs of the
 * row (for 2p: its even and its odd columns) as cyclic convolutions of length p - 1
 * (fft_len = p - 1), about half the row-pass work of BLUESTEIN.  SPLIT: for cols = n1 * n2 with
 * a length-n1 column kernel (n1 <= 64 with prime factors <= 31, or an odd prime <= 127)
 * and an on-chip row pass of width n2 (Rader's convolution of length n2 - 1 for an odd
 * prime n2 with a run-kernel shape, else DIRECT, or BLUESTEIN at a fused length): the
 * row DFT as n1 DFTs of length n2 on chip, then n1-point DFTs across them, through a
 * tile buffer of cols entries per row (m1 = n1, m2 = n2 in spfft_plan_info; FP32 form,
 * no fft_len / conv_rows / four-step options).  AUTO: DIRECT, else RADER, when possible;
 * else BLUESTEIN when its length has an on-chip shape; else SPLIT when possible; else
 * BLUESTEIN with the four-step kernels.  The transform computed is the same; only the
 * rounding differs. */
enum { SPFFT_TRANSFORM_AUTO = 0, SPFFT_TRANSFORM_BLUESTEIN = 1, SPFFT_TRANSFORM_DIRECT = 2,
       SPFFT_TRANSFORM_RADER = 3, SPFFT_TRANSFORM_SPLIT = 4 };

typedef struct spfft_plan_info {
    int struct_size;   /* sizeof(spfft_plan_info), set by the caller      */
    int rows, cols;
    int out_cols;      /* cols/2 + 1                                   */
    int pitch;         /* P, elements                                  */
    int fft_len;       /* M (SPLIT: the width-m2 pass's length: m2, m2 - 1 by Rader, or Bluestein's) */
    int m1, m2;        /* M = m1 * m2 (SPLIT: cols = m1 * m2)             */
    int packed;        /* 1 if CSC indices are stored as uint16        */
    int u_tile;
    int conv_rows;          /* rows per workgroup of the convolution kernel            */
    int conv_fast;          /* 1: register-codelet kernel, 0: generic LDS fallback     */
    int conv_stages;        /* radix stages of the length-m2 FFT                       */
    int conv_radix[8];      /* their radices                                           */
    int colpass_rows;       /* output rows per lane in the column pass                 */
    int colpass_full_table; /* 1: whole W_R table in LDS, 0: two-level table           */
    int precision;          /* SPFFT_PREC_* of the length-m1 sub-FFTs                   */
    size_t workspace_bytes; /* device bytes held by the plan (tables + tile buffer) */
    int row_kernel;         /* SPFFT_ROW_FOUR_STEP or SPFFT_ROW_FUSED                   */
    int colpass_paired;     /* 1 if the column pass pairs rows u and R/2 - u            */
    int direct;             /* 1: direct length-cols FFT (fft_len = cols), 0: Bluestein  */
    int rader;              /* 1: Rader convolutions of length fft_len = cols/2 - 1      */
    int row_transform;      /* SPFFT_TRANSFORM_BLUESTEIN, _DIRECT, _RADER or _SPLIT (resolved) */
    int t_rows;             /* four-step kernels: rows of T per sub-tile (0: fused)      */
    int row_threads;        /* on-chip row kernels: threads per workgroup (0: four-step) */
} spfft_plan_info;


typedef struct spfft_plan_s* spfft_plan;

/* Result of the reference-shaped one-shot entry points (mirrors SparseFFTResult,
 * include/sparse_fft.h:7-13).  d_output is owned by the caller: release it with
 * spfft_free_output().  ms = GPU time of the tile loop only (the reference's
 * timed region, src/sparse_fft_bluestein.cu:2322-2377); ms_e2e = GPU time of
 * the CSC build plus the tile loop. */
typedef struct spfft_result {
    int struct_size; /* sizeof(spfft_result), set by the caller */
    void* d_output;
    int output_cols; /* pitch P */
    float ms;
    float ms_e2e;
    size_t mem_bytes; /* device free-memory delta over the call */
    int fft_len;       /* the plan's M (Bluestein length, or cols / cols/2 - 1) */
    int row_transform; /* SPFFT_TRANSFORM_BLUESTEIN, _DIRECT or _RADER        */
} spfft_result;

/* Device view of the compact CSC built by the last execute / csc_build call
 * (the reference's CompactCSC after packing, include/sparse_fft_internal.cuh:
 * 193-222 and :146-151). */
typedef struct spfft_csc_view {
    int n_sparse_cols;
    int nnz;
    int packed;                /* 1: uint16 arrays, 0: int32 arrays  */
    const void* d_sparse_cols; /* n_sparse_cols entries               */
    const int* d_col_ptr;      /* n_sparse_cols + 1 entries           */
    const void* d_row_idx;     /* nnz entries                         */
    const int* d_col_start;    /* cols + 1 dense column starts        */
} spfft_csc_view;

void spfft_opts_default(spfft_opts* opts);
const char* spfft_last_error(void);
const char* spfft_version(void);

/* Planning: tables (chirp, B^, twiddles) in FP64 rounded to FP32, device
 * workspace for one tile. */
spfft_status spfft_plan_create(int rows, int cols, const spfft_opts* opts, spfft_plan* plan);
spfft_status spfft_plan_destroy(spfft_plan plan);
/* Shape planning only, on the host (no device needed): fills `info` (workspace_bytes = 0)
 * for the plan spfft_plan_create would build with a 64 KiB LDS budget. */
spfft_status spfft_plan_describe(int rows, int cols, const spfft_opts* opts, spfft_plan_info* info);
spfft_status spfft_plan_get_info(spfft_plan plan, spfft_plan_info* info);

/* Stage 1 only: device COO (int32 rows, cols) -> compact CSC held by the plan. */
spfft_status spfft_csc_build(spfft_plan plan, const int* d_row, const int* d_col, int nnz, void* stream);
/* Stage 1 with no device allocation.  No counterpart in the reference: it is for
 * callers that cannot call hipMalloc (code under hipGraph capture, tight serving
 * loops, harnesses that forbid allocation).  The same kernels and output as
 * spfft_csc_build + spfft_csc_get, into caller-owned device arrays:
 *   d_sparse_cols  cols entries (uint16 when rows, cols <= 65536, else int32); the
 *                  first d_info[1] are the active columns
 *   d_col_ptr      cols + 1 int32 (the first d_info[1] + 1 are written)
 *   d_row_idx      nnz entries (uint16 / int32 as d_sparse_cols)
 *   d_col_start    cols + 1 int32 dense column starts
 *   d_info         3 int32: [0] stored nonzeros, [1] active columns, [2] nonzero if an
 *                  index was outside [0, rows) x [0, cols) (such pairs are dropped)
 * The temporaries live in one file-scope device array of SPFFT_NOALLOC_SCRATCH_BYTES
 * per GPU, held until the process ends: cols <= SPFFT_NOALLOC_MAX_COLS and
 * nnz <= SPFFT_NOALLOC_MAX_NNZ, else SPFFT_ERR_INVALID_ARG.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (other host threads, or
 * other streams without an ordering between them) race.  Asynchronous on `stream`. */
#define SPFFT_NOALLOC_MAX_COLS 65536
#define SPFFT_NOALLOC_MAX_NNZ (1 << 20)
#define SPFFT_NOALLOC_SCRATCH_BYTES (18u << 20)
spfft_status spfft_csc_build_noalloc(int rows, int cols, const int* d_row, const int* d_col, int nnz,
                                     void* d_sparse_cols, int* d_col_ptr, void* d_row_idx, int* d_col_start,
                                     int* d_info, void* stream);
/* Synchronises `stream`, then fills `view` (and checks the input range flag). */
spfft_status spfft_csc_get(spfft_plan plan, void* stream, spfft_csc_view* view);
/* Stages 2 + 3 on the CSC of the last spfft_csc_build: writes d_out (pitch in elements). */
spfft_status spfft_rowpass(spfft_plan plan, void* d_out, int64_t out_pitch, void* stream);
/* All three stages, asynchronous on `stream`.  Inputs and output in device memory. */
spfft_status spfft_execute(spfft_plan plan, const int* d_row, const int* d_col, int nnz, void* d_out,
                           int64_t out_pitch, void* stream);
/* All three stages with the output streamed to host memory h_out (rows x out_pitch
 * complex64; pinned memory from hipHostMalloc lets the copies overlap the compute).
 * Only two chunks of 2 * u_tile output rows are held on the device, so outputs
 * larger than device memory can be produced: each tile's rows and mirror rows are
 * copied on a plan-owned copy stream while the next tile computes.  Asynchronous
 * on `stream`: synchronise it before reading h_out.  Replaces the streaming
 * variants sparse_fft_csc_stockham_streaming / sparse_fft_csc_bluestein_cufft_streaming
 * (reference include/sparse_fft.h:35-37, src/sparse_fft_bluestein.cu:912-1145). */
spfft_status spfft_execute_to_host(spfft_plan plan, const int* d_row, const int* d_col, int nnz,
                                   void* h_out, int64_t out_pitch, void* stream);
/* Stages 2 + 3 of spfft_execute_to_host on the CSC of the last spfft_csc_build. */
spfft_status spfft_rowpass_to_host(spfft_plan plan, void* h_out, int64_t out_pitch, void* stream);
/* Synchronises `stream` and reports SPFFT_ERR_BAD_INPUT if the last CSC build
 * saw an index outside the matrix. */
spfft_status spfft_check_input(spfft_plan plan, void* stream);

/* Per-kernel GPU time from HIP events recorded on the launch stream around
 * the CSC build and around each row-pass kernel of every tile (the role of the
 * reference's CUSPFFT_BREAKDOWN timing, src/sparse_fft_bluestein.cu:1752-1760).
 * Index 0 = CSC build, 1 = column pass, 2 = length-m1 sub-FFTs, 3 = convolution
 * rows (length-m2 FFT, B^ product, inverse), 4 = inverse length-m1 sub-FFTs +
 * finalize.  spfft_profile_read synchronises `stream`, returns the sums since
 * the last read and resets them. */
typedef struct spfft_profile {
    double ms[5];
    int launches[5];
} spfft_profile;
spfft_status spfft_profile_enable(spfft_plan plan, int on);
spfft_status spfft_profile_read(spfft_plan plan, void* stream, spfft_profile* out);

/* Reference-shaped synchronous entry points: host COO in, device output out.
 *   spfft_csc_bluestein_mixed_radix   replaces sparse_fft_csc_bluestein_mixed_radix
 *                                     (include/sparse_fft.h:57, src/sparse_fft_bluestein.cu:2256-2392)
 *   spfft_csc_bluestein_cufft_smooth  replaces sparse_fft_csc_bluestein_cufft_smooth
 *                                     (include/sparse_fft.h:70, src/sparse_fft_bluestein.cu:2557-2719)
 * Both always run the reference's algorithm: the Bluestein row pass at
 * M = next_3_smooth(2*cols-1) (mixed radix) or next_7_smooth(2*cols-1) (cuFFT
 * smooth), with u_tile rows per tile (spfft_reference_opts).  The direct and
 * Rader row transforms, which compute the same half spectrum with less work, are
 * reached through spfft_opts (row_transform AUTO) and spfft_run_host_coo.
 * u_tile <= 0 is rejected with SPFFT_ERR_INVALID_ARG ("CSC tile size must be
 * positive", as the reference throws at :2264-2265). */
/* The options of the reference-shaped entry points: length_rule = rule (3 or 7),
 * u_tile, row_transform = SPFFT_TRANSFORM_BLUESTEIN, the rest defaults. */
void spfft_reference_opts(int rule, int u_tile, spfft_opts* opts);
spfft_status spfft_csc_bluestein_mixed_radix(int rows, int cols, int nnz, const int* h_row,
                                             const int* h_col, int u_tile, spfft_result* result);
spfft_status spfft_csc_bluestein_cufft_smooth(int rows, int cols, int nnz, const int* h_row,
                                              const int* h_col, int u_tile, spfft_result* result);
/* General form of the two above (any opts). */
spfft_status spfft_run_host_coo(int rows, int cols, int nnz, const int* h_row, const int* h_col,
                                const spfft_opts* opts, spfft_result* result);
/* The streamed form of spfft_run_host_coo (spfft_execute_to_host): the output goes
 * to the caller's host buffer h_out (rows x out_pitch complex64, pinned for
 * overlap); result->d_output is NULL, result->output_cols = out_pitch, and ms
 * includes the device-to-host copies, as the reference's streaming drivers time
 * (src/sparse_fft_bluestein.cu:1035-1116). */
spfft_status spfft_run_host_coo_to_host(int rows, int cols, int nnz, const int* h_row, const int* h_col,
                                        const spfft_opts* opts, void* h_out, int64_t out_pitch,
                                        spfft_result* result);
spfft_status spfft_free_output(void* d_output);

/* ---- hipGraph form ------------------------------------------------------
 * The CSC build and the whole tile loop of a plan captured once into a hipGraph
 * and replayed with one launch (the reference's CUDA-graph variant,
 * src/sparse_fft_bluestein.cu:2402-2549).  Device pointers are baked into the
 * graph: replays read the same COO buffers and write the same output. */
typedef struct spfft_graph_s* spfft_graph;
spfft_status spfft_graph_capture(spfft_plan plan, const int* d_row, const int* d_col, int nnz, void* d_out,
                                 int64_t out_pitch, void* stream, spfft_graph* graph);
spfft_status spfft_graph_launch(spfft_graph graph, void* stream);
spfft_status spfft_graph_destroy(spfft_graph graph);
/* Replaces sparse_fft_csc_bluestein_mixed_radix_graph (include/sparse_fft.h:63): ms is
 * the GPU time of the single graph launch. */
spfft_status spfft_csc_bluestein_mixed_radix_graph(int rows, int cols, int nnz, const int* h_row,
                                                   const int* h_col, int u_tile, spfft_result* result);

/* ---- batched form --------------------------------------------------------
 * `count` matrices of uneven shape, each with its own plan (opts: `count`
 * entries, or NULL for defaults with one tile per matrix), run on `streams`
 * worker streams (longest first) that share one row-pass workspace per stream.
 * Inputs: all COO pairs concatenated in device memory, matrix i at
 * [coo_off[i], coo_off[i+1]) (host array of count+1).  Output: one arena; matrix
 * i at element offset offsets[i] with pitch pitches[i] (spfft_batch_layout).
 * nnz_hint (may be NULL) only steers the assignment to streams. */
typedef struct spfft_batch_s* spfft_batch;
spfft_status spfft_batch_create(int count, const int* rows, const int* cols, const int* nnz_hint,
                                const spfft_opts* opts, int streams, spfft_batch* batch);
spfft_status spfft_batch_destroy(spfft_batch batch);
spfft_status spfft_batch_layout(spfft_batch batch, int64_t* total_elems, int64_t* offsets, int* pitches,
                                spfft_plan_info* infos);
spfft_status spfft_batch_execute(spfft_batch batch, const int* d_row, const int* d_col,
                                 const int64_t* coo_off, void* d_out, void* stream);
/* Captures the fork / per-matrix work / join sequence of spfft_batch_execute into one graph. */
spfft_status spfft_batch_capture(spfft_batch batch, const int* d_row, const int* d_col,
                                 const int64_t* coo_off, void* d_out, spfft_graph* graph);
/* Synchronises `stream`; SPFFT_ERR_BAD_INPUT names the first matrix with an out-of-range index. */
spfft_status spfft_batch_check_input(spfft_batch batch, void* stream);
/* How the batch runs.  AUTO (the default): the matrices whose row pass is one on-chip
 * kernel (fused, direct or Rader plans with one tile; every matrix with rows, cols <=
 * 65536) -- the members -- run as one launch sequence for all of them: one CSC build of
 * their block-diagonal matrix, then per kernel shape one launch whose workgroups take
 * (matrix, block) items from a device table of the matrices' arguments.  The other
 * matrices keep their own launches on the worker streams, concurrently.  PER_MATRIX:
 * every matrix on its own launches on the worker streams.  Both compute the same output.
 * A mode change rebuilds the batch's shared buffers: graphs captured before it must not be
 * launched after it. */
enum { SPFFT_BATCH_AUTO = 0, SPFFT_BATCH_PER_MATRIX = 1 };
spfft_status spfft_batch_set_mode(spfft_batch batch, int mode);
/* The mode, and how many matrices run through the shared launches (members). */
spfft_status spfft_batch_get_mode(spfft_batch batch, int* mode, int* members);

/* ---- learned plan selector network ----------------------------------------
 * A D -> H -> H -> 1 ReLU MLP in FP32 with hand-written forward, backward and
 * Adam kernels.  Feature rows are row-major [n_rows x D] floats in device
 * memory.  Parameters: W1 [H x D], b1 [H], W2 [H x H], b2 [H], W3 [H], b3 [1]. */
typedef struct spfft_mlp_s* spfft_mlp;
spfft_status spfft_mlp_create(int in_dim, int hidden, uint64_t seed, spfft_mlp* mlp);
spfft_status spfft_mlp_destroy(spfft_mlp mlp);
int64_t spfft_mlp_num_params(spfft_mlp mlp);
spfft_status spfft_mlp_get_params(spfft_mlp mlp, float* h_params);
spfft_status spfft_mlp_set_params(spfft_mlp mlp, const float* h_params);
spfft_status spfft_mlp_predict(spfft_mlp mlp, const float* d_X, int n_rows, float* d_y, void* stream);
/* Loss 1/(2B) sum (yhat - y)^2 of one batch and its gradient (host, parameter layout). */
spfft_status spfft_mlp_loss_grad(spfft_mlp mlp, const float* d_X, const float* d_y, int batch, float* loss,
                                 float* h_grads, void* stream);
spfft_status spfft_mlp_train(spfft_mlp mlp, const float* d_X, const float* d_y, int n_rows, int epochs,
                             int batch, float lr, float weight_decay, uint64_t seed, float* last_loss,
                             void* stream);

/* Smooth lengths and the radix order of the reference planner
 * (src/sparse_fft_bluestein.cu:1975-2011). */
int spfft_next_3_smooth(int n);
int spfft_next_7_smooth(int n);
/* Writes the factor_3_smooth radix order into out (cap entries); returns the
 * count, or -1 if n is not 3-smooth. */
int spfft_factor_3_smooth(int n, int* out, int cap);
/* Bluestein lengths M (7-smooth) that have a fused row-kernel shape (SPFFT_ROW_FUSED), in
 * increasing order: writes up to cap of them to out and returns how many exist.
 * No counterpart in the reference (its row pass has no fused kernel). */
int spfft_fused_lengths(int* out, int cap);
/* Lengths cols with a direct transform (SPFFT_TRANSFORM_DIRECT: a single-batch fused
 * shape of length cols), increasing; same convention as spfft_fused_lengths.
 * Replaces the reference's power-of-two-only plain FFT paths for such widths
 * (README.md:653-673: Bluestein doubles the work of a power-of-two width). */
int spfft_direct_lengths(int* out, int cap);
/* Widths with a Rader row transform (SPFFT_TRANSFORM_RADER), increasing; same
 * convention: cols = 2p for p - 1 = 29 m (1858, 4178, 16706), and cols = p or 2p
 * from 1024 to 16384 for every prime p whose p - 1 is 7-smooth and has a kernel shape.
 * No counterpart in the reference, whose every non-power-of-two width goes through
 * Bluestein (src/sparse_fft_bluestein.cu:2327-2370). */
int spfft_rader_widths(int* out, int cap);

/* Seeded synthetic binary matrices (host memory, canonical COO: deduplicated,
 * sorted by (row, col), zero diagonal, symmetric).  Release with spfft_coo_free. */
typedef struct spfft_host_coo {
    int rows, cols, nnz;
    int* row_idx;
    int* col_idx;
} spfft_host_coo;

/* kind: "uniform" (p0 = number of distinct i<j pairs),
 *       "banded"  (p0 = target nnz, p1 = decay scale, p2 = band cutoff, p3 = off-band fraction),
 *       "chunglu" (p0 = target nnz, p1 = degree exponent, p2 = max expected degree, 0 = sqrt(target)). */
spfft_status spfft_coo_generate(const char* kind, int n, double p0, double p1, double p2, double p3,
                                uint64_t seed, spfft_host_coo* out);
void spfft_coo_free(spfft_host_coo* coo);
/* Matrix Market reader with the reference's semantics (src/mtx_reader.cu:7-77):
 * symmetric entries mirrored, 1-based -> 0-based, duplicates removed, sorted by (row, col). */
spfft_status spfft_read_mtx(const char* path, spfft_host_coo* out);

#ifdef __cplusplus
}
#endif

#endif /* SPFFT_C_H */
