/*
 * gpufsm_hip.h — C ABI of the MI355X (gfx950) HIP backend of gpufsm.
 *
 * Plain C types only (pointers, sizes, int status codes). Loaded from Python with ctypes
 * (gpufsm/backends/hip/_lib.py); any other FFI can bind the same symbols.
 *
 * What this replaces in the upstream project: the pybind11 entry points of the CUDA
 * extension, declared in `src/gpufsm/backends/cuda/native/include/api.hpp:23-104` and
 * bound in `src/gpufsm/backends/cuda/native/bindings.cu:7-39`. Those take numpy CSR arrays
 * and an int32 input batch, copy everything to the device on every call, run one kernel and
 * return `(flags, lens, kernel_ms)`. Here the automaton is uploaded once into an opaque
 * handle that stays resident in HBM, inputs are bytes with int64 offsets, and each run
 * either takes host buffers (synchronous, the drop-in form) or device buffers on a caller
 * stream (asynchronous, the resident form).
 *
 * Conventions
 *   - Every function returns GFSM_OK (0) or a positive GFSM_E* code; the message of the
 *     last failure on the calling thread is gfsm_last_error().
 *   - The caller owns every buffer it passes; the library never keeps a pointer to a
 *     caller buffer after the call returns (host forms) or after the stream reaches the
 *     end of the enqueued work (device forms). The library owns device memory of handles.
 *   - Results per input: flags[i] in {0,1} (accepted), lens[i] = match length
 *     (latch-first-match: bytes consumed up to and including the first byte after which an
 *     accepting state is active; 0 if the start closure accepts or nothing accepts).
 *   - Device forms are asynchronous on `stream` (a hipStream_t passed as void*; NULL means
 *     the null stream). Calls on one handle from several host threads must use distinct
 *     streams; handles are read-only after creation, so concurrent runs are safe.
 */
#ifndef GPUFSM_HIP_H
#define GPUFSM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GFSM_ABI_VERSION 4

/* status codes */
#define GFSM_OK 0
#define GFSM_EINVAL 1        /* bad argument (null pointer, out-of-range index, ...) */
#define GFSM_ENOMEM 2        /* device or host allocation failed */
#define GFSM_EUNSUPPORTED 3  /* technique cannot run this automaton / size */
#define GFSM_EDEVICE 4       /* a HIP runtime call failed */

/* NFA techniques (upstream names in parentheses are the CUDA kernels each one replaces) */
#define GFSM_NFA_AUTO 0       /* pick from the automaton's structure and batch size */
#define GFSM_NFA_WORKLIST 1   /* one wave64 per input over the compacted active list:
                                 successors ORed into an LDS bitmap, compacted by popcount
                                 and a wave prefix sum (worklist_warp / worklist_shared /
                                 worklist_compact); any NFA */
#define GFSM_NFA_BITPACKED 2  /* bit-parallel step (bitpacked / multistream / shared): for
                                 chain-structured NFAs the shift step over 4096-state tiles
                                 with class masks in LDS, for any other NFA a bitset per
                                 input in LDS updated from the byte's edge list (up to 65,536
                                 states; epsilon closure folded on the host) */
#define GFSM_NFA_CHAIN 3      /* sparse chain step: persistent waves over (input, chunk)
                                 items, active slots listed in LDS and compacted by ballot;
                                 long inputs cut into overlapping chunks when the start state
                                 loops on every byte; chain-structured NFAs only */
#define GFSM_NFA_REGISTER 4   /* one lane per input, the state set in registers, up to 512
                                 states (worklist_multistream, the register worklist) */

/* DFA techniques */
#define GFSM_DFA_AUTO 0
#define GFSM_DFA_GATHER 1     /* one lane per input, table in global memory (dfa_kernel) */
#define GFSM_DFA_LDS 2        /* one lane per input, uint16 table staged in LDS */
#define GFSM_DFA_SPLIT 3      /* exact in-input parallel walk: fixed segments of the batch,
                                 speculative entries repaired by checkpoint merging */

typedef struct gfsm_nfa gfsm_nfa;
typedef struct gfsm_dfa gfsm_dfa;

/* Tuning of one run (nullable everywhere; a zero field means "the library's choice"). No
 * upstream counterpart: upstream kernels have fixed launch shapes. */
#define GFSM_TUNE_CHAIN_FIRST 1  /* chain: no byte-pair lists (first-position lists only) */
#define GFSM_TUNE_CHAIN_WAVE64 2 /* chain: one work item per wave in the register tier (64
                                    listed states) instead of two (32 each) */
#define GFSM_TUNE_CHAIN_TRIPLE 4 /* chain: the lane tier (one work item per lane) over
                                    byte-triple lists of the states entered at depth 3, for
                                    pattern unions whose start state loops on every byte and
                                    whose slots fit 16 bits; ignored for other automata. The
                                    lists (64 MiB of headers plus the lists) are built and
                                    uploaded by the first run with this flag, which therefore
                                    must not be inside stream capture */
typedef struct {
    int32_t chunk_bytes;      /* chain: bytes per work item when inputs are cut (>= 64) */
    int32_t segment_bytes;    /* dfa split: segment length, a power of two in [256, 1 MiB] */
    int32_t flags;            /* GFSM_TUNE_* bits */
    int32_t max_input_bytes;  /* longest input of the batch, 0 = not known. The DFA auto
                                 rule walks one lane per input only when it is known and
                                 short against the batch (a lane walks its input alone; one
                                 long input among many short ones would serialise the run);
                                 the host forms compute it themselves */
} gfsm_tuning;

typedef struct {
    int32_t num_states;
    int32_t start_state;
    int64_t num_sym_edges;
    int64_t num_eps_edges;
    int32_t start_accepts;    /* start closure holds an accepting state */
    int32_t chain_ok;         /* GFSM_NFA_BITPACKED applies */
    int32_t num_tiles;        /* 4096-state tiles of the bitpacked layout (0 if !chain_ok) */
    int32_t start_always;     /* start state has a wildcard self-loop */
    int32_t max_chain;        /* longest chain (bytes a match can span) if chain_ok */
    int32_t general_ok;       /* the epsilon-folded layout of the general techniques exists
                                 (automata that are not chain-structured, and any automaton
                                 of at most 512 states, when the fold stays under its cap) */
    int64_t device_bytes;     /* HBM held by the handle (with the chain byte-triple lists
                                 once a GFSM_TUNE_CHAIN_TRIPLE run has built them) */
} gfsm_nfa_props;

typedef struct {
    int32_t num_states;
    int32_t start_state;
    int32_t start_accepts;
    int32_t entry_bytes;      /* 2 (uint16 table) or 4 (int32 table) */
    int32_t lds_ok;           /* GFSM_DFA_LDS applies */
    int32_t pad;
    int64_t device_bytes;
} gfsm_dfa_props;

/* Library */
int gfsm_abi_version(void);
const char *gfsm_last_error(void);
/* Name of device `ordinal`'s architecture (e.g. "gfx950") into buf. */
int gfsm_device_arch(int32_t ordinal, char *buf, int32_t buflen);

/* NFA handles. Replaces the per-call CSR upload of every run_* in api.hpp:39-97
 * (dev_copy of sym_row_ptr/sym_targets/sym_symbols/eps_row_ptr/eps_targets/accept). The
 * arrays follow `src/gpufsm/core/nfa.py:43-65`: labels 0..255 or 256 (= ANY_SYMBOL). */
int gfsm_nfa_create(int32_t num_states, int32_t start_state, const uint8_t *accept,
                    const int32_t *sym_row_ptr, const int32_t *sym_targets,
                    const int32_t *sym_symbols, const int32_t *eps_row_ptr,
                    const int32_t *eps_targets, gfsm_nfa **out);
int gfsm_nfa_destroy(gfsm_nfa *nfa);
int gfsm_nfa_query(const gfsm_nfa *nfa, gfsm_nfa_props *out);
/* Structural analysis on the host only (no device needed, nothing uploaded): fills the
 * same properties gfsm_nfa_query reports (device_bytes = what a handle would hold). No
 * upstream counterpart; technique selection uses it before committing device memory. */
int gfsm_nfa_analyze(int32_t num_states, int32_t start_state, const uint8_t *accept,
                     const int32_t *sym_row_ptr, const int32_t *sym_targets,
                     const int32_t *sym_symbols, const int32_t *eps_row_ptr,
                     const int32_t *eps_targets, gfsm_nfa_props *out);

/* NFA batch, host buffers, synchronous. Replaces run_worklist / run_worklist_warp /
 * run_multistream / ... (api.hpp:47-97): same inputs and outputs, bytes instead of int32
 * symbols and int64 offsets[num_strings + 1]. kernel_ms (nullable) is the device time of
 * the launch(es), measured with HIP events as the upstream entry points do. Each input is
 * at most 2^31 - 1 bytes (match_len is int32); the batch itself may exceed 2 GiB. */
int gfsm_nfa_run_host(gfsm_nfa *nfa, int32_t technique, const uint8_t *data,
                      const int64_t *offsets, int32_t num_strings, int32_t *flags,
                      int32_t *lens, float *kernel_ms);

/* NFA batch, device buffers, asynchronous on `stream`. No upstream counterpart (upstream
 * always copies from the host); it is the resident form the batched and served paths use.
 * total_bytes must equal d_offsets[num_strings] (the host knows it; reading it back would
 * cost a synchronisation): techniques that cut inputs into segments size scratch from it.
 * The offsets are not read back for checking: they must start at 0, never decrease, and
 * describe inputs of at most 2^31 - 1 bytes, as for the host form. */
int gfsm_nfa_run_device(gfsm_nfa *nfa, int32_t technique, const uint8_t *d_data,
                        const int64_t *d_offsets, int32_t num_strings, int64_t total_bytes,
                        int32_t *d_flags, int32_t *d_lens, void *stream);

/* Device scratch (bytes) that gfsm_nfa_run_device_ws needs for this technique and batch
 * shape. No upstream counterpart. */
int gfsm_nfa_workspace_size(gfsm_nfa *nfa, int32_t technique, int32_t num_strings,
                            int64_t total_bytes, const gfsm_tuning *tuning,
                            int64_t *out_bytes);

/* gfsm_nfa_run_device with tuning and caller-provided device scratch (at least
 * gfsm_nfa_workspace_size bytes, 256-byte aligned): enqueues kernels only, no allocation
 * and no memset, so the call can be captured into a HIP graph. No upstream counterpart
 * (it is the persistent/captured run_batch of the batched path). */
int gfsm_nfa_run_device_ws(gfsm_nfa *nfa, int32_t technique, const gfsm_tuning *tuning,
                           const uint8_t *d_data, const int64_t *d_offsets,
                           int32_t num_strings, int64_t total_bytes, int32_t *d_flags,
                           int32_t *d_lens, void *d_workspace, int64_t workspace_bytes,
                           void *stream);

/* The concrete technique GFSM_NFA_AUTO (or `technique`) resolves to for this batch shape
 * and tuning (nullable); GFSM_EUNSUPPORTED if the technique cannot run the automaton. No
 * upstream counterpart. */
int gfsm_nfa_resolve(gfsm_nfa *nfa, int32_t technique, const gfsm_tuning *tuning,
                     int32_t num_strings, int64_t total_bytes, int32_t *out_technique);

/* DFA handles. Replaces the per-call trans/accept upload of run_dfa (api.hpp:101-104);
 * layout of `src/gpufsm/core/dfa.py:25-37` (trans[s * 256 + c]). */
int gfsm_dfa_create(int32_t num_states, int32_t start_state, const uint8_t *accept,
                    const int32_t *trans, gfsm_dfa **out);
int gfsm_dfa_destroy(gfsm_dfa *dfa);
int gfsm_dfa_query(const gfsm_dfa *dfa, gfsm_dfa_props *out);

/* DFA batch, host buffers, synchronous. Replaces run_dfa (`native/dfa.cu:32-71`). */
int gfsm_dfa_run_host(gfsm_dfa *dfa, int32_t technique, const uint8_t *data,
                      const int64_t *offsets, int32_t num_strings, int32_t *flags,
                      int32_t *lens, float *kernel_ms);

/* DFA batch, device buffers, asynchronous on `stream` (resident form); total_bytes as for
 * gfsm_nfa_run_device. */
int gfsm_dfa_run_device(gfsm_dfa *dfa, int32_t technique, const uint8_t *d_data,
                        const int64_t *d_offsets, int32_t num_strings, int64_t total_bytes,
                        int32_t *d_flags, int32_t *d_lens, void *stream);

/* DFA counterparts of gfsm_nfa_workspace_size / gfsm_nfa_run_device_ws /
 * gfsm_nfa_resolve. No upstream counterpart. */
int gfsm_dfa_workspace_size(gfsm_dfa *dfa, int32_t technique, int32_t num_strings,
                            int64_t total_bytes, const gfsm_tuning *tuning,
                            int64_t *out_bytes);
int gfsm_dfa_run_device_ws(gfsm_dfa *dfa, int32_t technique, const gfsm_tuning *tuning,
                           const uint8_t *d_data, const int64_t *d_offsets,
                           int32_t num_strings, int64_t total_bytes, int32_t *d_flags,
                           int32_t *d_lens, void *d_workspace, int64_t workspace_bytes,
                           void *stream);
int gfsm_dfa_resolve(gfsm_dfa *dfa, int32_t technique, const gfsm_tuning *tuning,
                     int32_t num_strings, int64_t total_bytes, int32_t *out_technique);

/* Allocation-free DFA run. No counterpart in the upstream project: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same inputs, outputs and asynchrony as
 * gfsm_dfa_run_device_ws, but the technique's scratch (the split walk's segment state) is a
 * library array of GFSM_STATIC_SCRATCH_BYTES in each GPU's memory, held until the process
 * ends; the call makes no device allocation (its first call on a device looks up that
 * device's array). A run needing more scratch returns GFSM_EUNSUPPORTED before enqueuing
 * anything (gfsm_dfa_workspace_size gives the need; about 19 MiB for 4096 x 64 KiB on a
 * 256-state table). Every call on one GPU shares the array: calls whose work overlaps on
 * one GPU (other streams, other threads) race and give wrong results; serialise them. */
#define GFSM_STATIC_SCRATCH_BYTES (32 << 20)
int gfsm_dfa_run_device_static(gfsm_dfa *dfa, int32_t technique, const gfsm_tuning *tuning,
                               const uint8_t *d_data, const int64_t *d_offsets,
                               int32_t num_strings, int64_t total_bytes, int32_t *d_flags,
                               int32_t *d_lens, void *stream);

/* ---- Serving loop ----------------------------------------------------------------------
 * Continuous batching of match requests over device-resident automata, run by a native
 * scheduler thread (no interpreter lock on the request path). No upstream counterpart:
 * upstream has no serving path; the north-star serving loop of this backend is this API.
 *
 *   gfsm_server_create      slots (HIP streams with pinned staging, device buffers and a
 *                           workspace each) on the current device;
 *   gfsm_server_add_nfa/dfa register resident handles (the caller keeps them alive until
 *                           gfsm_server_destroy returns) before gfsm_server_start;
 *   gfsm_server_submit      copies the request's bytes and queues it; returns its id
 *                           (gfsm_server_submit_many: many requests in one call);
 *   gfsm_server_wait        waits for requests (timeout_s < 0: no limit, 0: poll) and
 *                           collects the finished ones: each id is collected exactly once;
 *   gfsm_server_stop        drain = 1 finishes every queued request first, 0 fails the ones
 *                           not yet launched (status GFSM_SERVE_ERROR);
 *   gfsm_server_destroy     stops without draining and frees the server (not the handles).
 *
 * Admission: a request is admitted while the resident automata plus the bytes of every
 * admitted, unfinished request stay within budget_bytes; one longer than max_batch_bytes or
 * than the budget left by the automata is rejected (GFSM_SERVE_REJECTED); one waiting
 * longer than queue_timeout_s (> 0) ends with GFSM_SERVE_TIMEOUT. A launch takes the oldest
 * admitted request whose automaton has fewer than max_inflight_per_automaton launches
 * running, plus the later admitted requests of that automaton that fit one slot. A device
 * error stops the loop: pending requests end with GFSM_SERVE_ERROR, later submits fail,
 * and gfsm_server_error holds the message. Times are steady-clock seconds
 * (CLOCK_MONOTONIC). Every entry point is thread-safe, except that gfsm_server_destroy must
 * not overlap any other call on the same server. */
#define GFSM_SERVE_PENDING 0
#define GFSM_SERVE_OK 1
#define GFSM_SERVE_REJECTED 2
#define GFSM_SERVE_TIMEOUT 3
#define GFSM_SERVE_ERROR 4

typedef struct gfsm_server gfsm_server;

typedef struct {              /* a zero field means the default */
    int32_t slots;            /* launch slots (streams), default 4, at most 64 */
    int32_t max_batch_count;  /* requests per launch, default 4096 */
    int64_t max_batch_bytes;  /* bytes per launch and longest request, default 16 MiB */
    int64_t budget_bytes;     /* HBM admission budget, default 32 GiB */
    int32_t max_inflight_per_automaton;  /* default 2 */
    int32_t poll_us;          /* completion polling interval while launches run, default 20 */
    double queue_timeout_s;   /* <= 0: no timeout */
} gfsm_server_config;

typedef struct {
    int32_t status;           /* GFSM_SERVE_* */
    int32_t accepted;
    int32_t match_len;
    int32_t automaton;
    int64_t bytes;            /* request length */
    double arrival_s, done_s; /* submit and completion times */
} gfsm_server_result;

typedef struct {
    int64_t submitted, completed, rejected, timed_out, errors, launches;
    int64_t credited_bytes;   /* bytes of the requests completed */
    int64_t waiting, admitted, in_flight_bytes, resident_bytes;
    int32_t failed, pad;
} gfsm_server_stats;

int gfsm_server_create(const gfsm_server_config *config, gfsm_server **out);
int gfsm_server_add_nfa(gfsm_server *server, gfsm_nfa *nfa, int32_t technique,
                        const gfsm_tuning *tuning, int32_t *out_index);
int gfsm_server_add_dfa(gfsm_server *server, gfsm_dfa *dfa, int32_t technique,
                        const gfsm_tuning *tuning, int32_t *out_index);
int gfsm_server_start(gfsm_server *server);
int gfsm_server_submit(gfsm_server *server, int32_t automaton, const uint8_t *data,
                       int64_t len, int64_t *out_id);
/* n requests for one automaton at once: request i is data[offsets[i], offsets[i + 1]);
 * ids in order into out_ids[n]; the requests are queued together, in order, with one
 * arrival time. The library copies the whole run once (with several threads when it is
 * large), unless flags holds GFSM_SERVE_NOCOPY: then it reads the caller's buffer when it
 * stages each launch (a buffer in pinned host memory is copied to the device directly,
 * without staging), and the buffer must stay valid and unchanged until every one of the
 * n requests has finished. */
#define GFSM_SERVE_NOCOPY 1
int gfsm_server_submit_many(gfsm_server *server, int32_t automaton, const uint8_t *data,
                            const int64_t *offsets, int32_t n, int32_t flags,
                            int64_t *out_ids);
int gfsm_server_wait(gfsm_server *server, const int64_t *ids, int64_t n, double timeout_s,
                     gfsm_server_result *out);
int gfsm_server_query(gfsm_server *server, gfsm_server_stats *out);
int gfsm_server_error(gfsm_server *server, char *buf, int32_t buflen);
int gfsm_server_stop(gfsm_server *server, int32_t drain);
int gfsm_server_destroy(gfsm_server *server);

#ifdef __cplusplus
}
#endif

#endif /* GPUFSM_HIP_H */
