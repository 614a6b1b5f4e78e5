/*
 * CPU oracle, C restatement. TEST INFRASTRUCTURE ONLY: it is the checker the tests and the
 * benchmark's CPU-baseline leg compare the HIP kernels against; the product never links it.
 *
 * Semantics restated from the upstream oracle `src/gpufsm/reference.py`:
 *   oracle_nfa_*  follows `simulate`      (reference.py:36-65) with the epsilon closure of
 *                 `epsilon_closure`       (reference.py:20-33);
 *   oracle_dfa_*  follows `simulate_dfa`  (reference.py:68-81).
 * Latch-first-match: an accepting start closure gives (1, 0); otherwise the first byte
 * index i whose post-step closure holds an accepting state gives (1, i + 1); an empty
 * successor set, or the end of input, gives (0, 0).
 *
 * Additionally reports "examined" bytes per input: match_len on accept, the index after
 * the byte that emptied the set, or the input length.
 *
 * Two NFA modes with identical verdicts:
 *   scan    each active state's edge row is scanned in full, as reference.py:54-58 does;
 *   indexed rows are pre-sorted by label so a step binary-searches the byte's edge run
 *           (plus the wildcard run). Same set of successors, far fewer comparisons on
 *           automata whose start state has thousands of edges.
 * Batches run on a pthread pool that takes inputs from a shared atomic counter.
 */
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ANY_SYMBOL 256

typedef struct {
    int32_t n, start;
    const uint8_t *accept;
    const int32_t *srp, *st, *ss, *erp, *et;
    /* indexed mode: per-state edges sorted by label (ANY = 256 sorts last) */
    int32_t *sorted_t, *sorted_s;
} nfa_t;

typedef struct {
    int32_t *mark;   /* generation stamp per state: dedup without clearing */
    int32_t gen;
    int32_t *cur, *nxt, *stack;
    int32_t ncur, nnxt;
} work_t;

static int work_init(work_t *w, int32_t n) {
    w->mark = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    w->cur = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    w->nxt = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    w->stack = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    w->gen = 0;
    return (w->mark && w->cur && w->nxt && w->stack) ? 0 : -1;
}

static void work_free(work_t *w) {
    free(w->mark); free(w->cur); free(w->nxt); free(w->stack);
}

static int32_t next_gen(work_t *w, int32_t n) {
    if (w->gen == INT32_MAX) {
        memset(w->mark, 0, (size_t)n * sizeof(int32_t));
        w->gen = 0;
    }
    return ++w->gen;
}

/* Extend list[0..*cnt) (all stamped `gen`) with its epsilon closure, depth-first. */
static void closure(const nfa_t *a, work_t *w, int32_t *list, int32_t *cnt, int32_t gen) {
    int32_t top = 0;
    for (int32_t i = 0; i < *cnt; ++i) w->stack[top++] = list[i];
    while (top) {
        int32_t s = w->stack[--top];
        for (int32_t k = a->erp[s]; k < a->erp[s + 1]; ++k) {
            int32_t t = a->et[k];
            if (w->mark[t] != gen) {
                w->mark[t] = gen;
                list[(*cnt)++] = t;
                w->stack[top++] = t;
            }
        }
    }
}

static int any_accept(const nfa_t *a, const int32_t *list, int32_t cnt) {
    for (int32_t i = 0; i < cnt; ++i)
        if (a->accept[list[i]]) return 1;
    return 0;
}

/* First index in [lo, hi) of sorted labels whose value is >= key. */
static int32_t lower(const int32_t *lab, int32_t lo, int32_t hi, int32_t key) {
    while (lo < hi) {
        int32_t mid = lo + ((hi - lo) >> 1);
        if (lab[mid] < key) lo = mid + 1; else hi = mid;
    }
    return lo;
}

static void nfa_one(const nfa_t *a, work_t *w, int indexed, const uint8_t *in, int64_t len,
                    int32_t *flag, int32_t *mlen, int64_t *examined) {
    int32_t gen = next_gen(w, a->n);
    w->ncur = 0;
    w->cur[w->ncur++] = a->start;
    w->mark[a->start] = gen;
    closure(a, w, w->cur, &w->ncur, gen);
    if (any_accept(a, w->cur, w->ncur)) { *flag = 1; *mlen = 0; *examined = 0; return; }
    for (int64_t i = 0; i < len; ++i) {
        int32_t b = in[i];
        gen = next_gen(w, a->n);
        w->nnxt = 0;
        for (int32_t j = 0; j < w->ncur; ++j) {
            int32_t s = w->cur[j];
            int32_t lo = a->srp[s], hi = a->srp[s + 1];
            if (!indexed) {
                for (int32_t k = lo; k < hi; ++k) {
                    int32_t sym = a->ss[k];
                    if (sym == b || sym == ANY_SYMBOL) {
                        int32_t t = a->st[k];
                        if (w->mark[t] != gen) { w->mark[t] = gen; w->nxt[w->nnxt++] = t; }
                    }
                }
            } else {
                for (int32_t k = lower(a->sorted_s, lo, hi, b); k < hi && a->sorted_s[k] == b; ++k) {
                    int32_t t = a->sorted_t[k];
                    if (w->mark[t] != gen) { w->mark[t] = gen; w->nxt[w->nnxt++] = t; }
                }
                for (int32_t k = lower(a->sorted_s, lo, hi, ANY_SYMBOL); k < hi; ++k) {
                    int32_t t = a->sorted_t[k];
                    if (w->mark[t] != gen) { w->mark[t] = gen; w->nxt[w->nnxt++] = t; }
                }
            }
        }
        if (w->nnxt == 0) { *flag = 0; *mlen = 0; *examined = i + 1; return; }
        closure(a, w, w->nxt, &w->nnxt, gen);
        int32_t *tmp = w->cur; w->cur = w->nxt; w->nxt = tmp; w->ncur = w->nnxt;
        if (any_accept(a, w->cur, w->ncur)) {
            *flag = 1; *mlen = (int32_t)(i + 1); *examined = i + 1; return;
        }
    }
    *flag = 0; *mlen = 0; *examined = len;
}

typedef struct {
    const nfa_t *nfa;
    int indexed;
    /* dfa */
    int32_t dn, dstart;
    const uint8_t *daccept;
    const int32_t *trans;
    /* batch */
    const uint8_t *data;
    const int64_t *offsets;
    int32_t count;
    int32_t *flags, *lens;
    int64_t *examined;
    atomic_int next;
    atomic_int failed;
} job_t;

static void dfa_one(const job_t *j, const uint8_t *in, int64_t len, int32_t *flag, int32_t *mlen,
                    int64_t *examined) {
    int32_t cur = j->dstart;
    if (j->daccept[cur]) { *flag = 1; *mlen = 0; *examined = 0; return; }
    for (int64_t i = 0; i < len; ++i) {
        cur = j->trans[(int64_t)cur * 256 + in[i]];
        if (j->daccept[cur]) { *flag = 1; *mlen = (int32_t)(i + 1); *examined = i + 1; return; }
    }
    *flag = 0; *mlen = 0; *examined = len;
}

static void *worker(void *arg) {
    job_t *j = (job_t *)arg;
    work_t w;
    int have_work = 0;
    if (j->nfa) {
        if (work_init(&w, j->nfa->n) != 0) { atomic_store(&j->failed, 1); work_free(&w); return NULL; }
        have_work = 1;
    }
    for (;;) {
        int i = atomic_fetch_add(&j->next, 1);
        if (i >= j->count) break;
        const uint8_t *in = j->data + j->offsets[i];
        int64_t len = j->offsets[i + 1] - j->offsets[i];
        int64_t ex = 0;
        if (j->nfa) nfa_one(j->nfa, &w, j->indexed, in, len, &j->flags[i], &j->lens[i], &ex);
        else dfa_one(j, in, len, &j->flags[i], &j->lens[i], &ex);
        if (j->examined) j->examined[i] = ex;
    }
    if (have_work) work_free(&w);
    return NULL;
}

static int run_pool(job_t *j, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    atomic_store(&j->next, 0);
    atomic_store(&j->failed, 0);
    pthread_t tid[256];
    int started = 0;
    for (int t = 1; t < nthreads; ++t) {
        if (pthread_create(&tid[started], NULL, worker, j) == 0) ++started;
    }
    worker(j);
    for (int t = 0; t < started; ++t) pthread_join(tid[t], NULL);
    return atomic_load(&j->failed) ? -1 : 0;
}

/* Sort each state's row by label (stable over targets is not needed: sets are compared). */
static int build_index(nfa_t *a) {
    int32_t nnz = a->srp[a->n];
    a->sorted_t = (int32_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    a->sorted_s = (int32_t *)malloc((size_t)(nnz ? nnz : 1) * sizeof(int32_t));
    if (!a->sorted_t || !a->sorted_s) return -1;
    int32_t count[258];
    for (int32_t s = 0; s < a->n; ++s) {
        int32_t lo = a->srp[s], hi = a->srp[s + 1];
        memset(count, 0, sizeof(count));
        for (int32_t k = lo; k < hi; ++k) count[a->ss[k] + 1]++;
        for (int c = 0; c < 257; ++c) count[c + 1] += count[c];
        for (int32_t k = lo; k < hi; ++k) {
            int32_t pos = lo + count[a->ss[k]]++;
            a->sorted_s[pos] = a->ss[k];
            a->sorted_t[pos] = a->st[k];
        }
    }
    return 0;
}

int oracle_nfa_batch(int32_t num_states, int32_t start_state, const uint8_t *accept,
                     const int32_t *sym_row_ptr, const int32_t *sym_targets,
                     const int32_t *sym_symbols, const int32_t *eps_row_ptr,
                     const int32_t *eps_targets, const uint8_t *data, const int64_t *offsets,
                     int32_t num_strings, int32_t *flags, int32_t *lens, int64_t *examined,
                     int32_t indexed, int32_t nthreads) {
    if (num_states <= 0 || start_state < 0 || start_state >= num_states || num_strings < 0)
        return -2;
    nfa_t a = {num_states, start_state, accept, sym_row_ptr, sym_targets, sym_symbols,
               eps_row_ptr, eps_targets, NULL, NULL};
    if (indexed && build_index(&a) != 0) { free(a.sorted_t); free(a.sorted_s); return -1; }
    job_t j;
    memset(&j, 0, sizeof(j));
    j.nfa = &a; j.indexed = indexed; j.data = data; j.offsets = offsets; j.count = num_strings;
    j.flags = flags; j.lens = lens; j.examined = examined;
    int rc = run_pool(&j, nthreads);
    free(a.sorted_t); free(a.sorted_s);
    return rc;
}

int oracle_dfa_batch(int32_t num_states, int32_t start_state, const uint8_t *accept,
                     const int32_t *trans, const uint8_t *data, const int64_t *offsets,
                     int32_t num_strings, int32_t *flags, int32_t *lens, int64_t *examined,
                     int32_t nthreads) {
    if (num_states <= 0 || start_state < 0 || start_state >= num_states || num_strings < 0)
        return -2;
    job_t j;
    memset(&j, 0, sizeof(j));
    j.dn = num_states; j.dstart = start_state; j.daccept = accept; j.trans = trans;
    j.data = data; j.offsets = offsets; j.count = num_strings;
    j.flags = flags; j.lens = lens; j.examined = examined;
    return run_pool(&j, nthreads);
}
