// Scheduler of the CPU build of the HIP kernels (test infrastructure).
//
// A launch runs its blocks one after another.  A kernel without barriers or
// wave exchanges runs as a plain loop over its threads; any other kernel runs
// each thread of a block as a fiber on its own stack.  A fiber runs until it
// reaches a block barrier or a wave exchange and yields; a barrier opens when
// every live thread of the block waits at it, a wave exchange when every live
// lane of the wave does.  Fibers are started once (ucontext) and switched with
// _setjmp/_longjmp, which save no signal mask and so make no system call.
#include "hip/hip_runtime.h"

#include <csetjmp>
#include <cstring>

dim3e threadIdx, blockIdx, blockDim, gridDim;

namespace emu {

bool needs_sync(const char *name) {
    while (*name == '(') ++name;   // a kernel name in parentheses (template arguments with commas)
    static const char *const sync[] = {"k_bounds_partial", "k_bounds_final", "k_count", "k_spu_nodes",
                                       "k_sp_update",      "k_deriv_nodes",  "k_scan_local", "k_scan_sums",
                                       "k_max_disp",       "k_gemm_bf16",    "k_mlp_loss",   "k_mlp_colsum",
                                       "k_surrogate_apply", "k_xsph",   "k_permute", "k_av_nodes",
                                       "k_cell_count",     "k_cell_scatter"};
    for (const char *s : sync)
        if (strncmp(name, s, strlen(s)) == 0) return true;   // k_count also covers k_count_lists
    return false;
}

unsigned char *lds_base = nullptr;
size_t lds_bytes = 0;

bool capturing = false;
struct Launch {
    dim3 grid;
    unsigned block;
    size_t lds;
    std::function<void()> body;
    bool sync;
};
struct Graph {
    std::vector<Launch> launches;
};
static Graph *recording = nullptr;

void check_not_capturing(const char *what) {
    if (capturing) {
        fprintf(stderr, "emu: %s while capturing a graph\n", what);
        abort();
    }
}

Graph *end_capture() {
    capturing = false;
    Graph *g = recording ? recording : new Graph();
    recording = nullptr;
    return g;
}

void replay(const Graph *g) {
    for (const Launch &l : g->launches) launch(l.grid, l.block, l.lds, l.body, l.sync);
}

void drop(Graph *g) { delete g; }

namespace {
enum State { RUN, AT_BLOCK, AT_WAVE, DONE };
struct Fiber {
    ucontext_t ctx;
    jmp_buf jb;
    State st = RUN;
    bool started = false;
    char *stack = nullptr;
};
constexpr size_t STACK = 512 * 1024;
constexpr int MAX_BLOCK = 1024;
Fiber fibers[MAX_BLOCK];
ucontext_t sched_ctx;
jmp_buf sched_jb;
int cur = -1;
bool fiber_mode = false;
const std::function<void()> *body_fn = nullptr;
unsigned char wave_buf[MAX_BLOCK / 64][64][16];
int or_acc = 0;

void yield_as(State s) {
    if (!fiber_mode) {
        fprintf(stderr, "emu: barrier or wave exchange in a kernel run without fibers\n");
        abort();
    }
    fibers[cur].st = s;
    if (!_setjmp(fibers[cur].jb)) _longjmp(sched_jb, 1);
}

void trampoline() {
    for (;;) {
        (*body_fn)();
        fibers[cur].st = DONE;
        if (!_setjmp(fibers[cur].jb)) _longjmp(sched_jb, 1);
    }
}

// Open whichever wave exchanges or block barrier every live thread waits at.
bool release(unsigned block) {
    bool any = false;
    for (unsigned w = 0; w * 64 < block; ++w) {
        int live = 0, waiting = 0;
        const unsigned hi = w * 64 + 64 < block ? w * 64 + 64 : block;
        for (unsigned l = w * 64; l < hi; ++l) {
            if (fibers[l].st == DONE) continue;
            ++live;
            waiting += fibers[l].st == AT_WAVE;
        }
        if (live && waiting == live) {
            for (unsigned l = w * 64; l < hi; ++l)
                if (fibers[l].st == AT_WAVE) fibers[l].st = RUN;
            any = true;
        }
    }
    if (any) return true;
    int live = 0, at = 0;
    for (unsigned t = 0; t < block; ++t) {
        if (fibers[t].st == DONE) continue;
        ++live;
        at += fibers[t].st == AT_BLOCK;
    }
    if (live == 0) return false;
    if (at != live) {
        fprintf(stderr, "emu: threads of block %u wait at different barriers\n", (unsigned)blockIdx.x);
        abort();
    }
    for (unsigned t = 0; t < block; ++t)
        if (fibers[t].st == AT_BLOCK) fibers[t].st = RUN;
    return true;
}
}  // namespace

void sync_block() { yield_as(AT_BLOCK); }

int sync_block_or(int p) {
    yield_as(AT_BLOCK);   // every earlier reader of the result is done
    if (threadIdx.x == 0) or_acc = 0;
    yield_as(AT_BLOCK);
    if (p) or_acc = 1;
    yield_as(AT_BLOCK);
    const int r = or_acc;
    yield_as(AT_BLOCK);
    return r;
}

void wave_put(const void *v, int bytes) {
    if (v) memcpy(wave_buf[threadIdx.x >> 6][threadIdx.x & 63], v, bytes);
    yield_as(AT_WAVE);
}

void wave_get(int lane, void *v, int bytes) { memcpy(v, wave_buf[threadIdx.x >> 6][lane], bytes); }

void launch(dim3 grid, unsigned block, size_t lds, const std::function<void()> &body, bool sync) {
    if (capturing) {
        if (!recording) recording = new Graph();
        recording->launches.push_back(Launch{grid, block, lds, body, sync});
        return;
    }
    if (block > MAX_BLOCK) {
        fprintf(stderr, "emu: block of %u threads\n", block);
        abort();
    }
    // exactly the requested dynamic LDS, so that an access past it is a heap overflow
    std::vector<unsigned char> ldsmem(lds ? lds : 1);
    lds_base = ldsmem.data();
    lds_bytes = lds;
    gridDim.x = grid.x;
    gridDim.y = grid.y;
    gridDim.z = grid.z;
    blockDim.x = block;
    body_fn = &body;
    fiber_mode = sync;
    const unsigned long long nblocks = (unsigned long long)grid.x * grid.y * grid.z;
    for (unsigned long long b = 0; b < nblocks; ++b) {
        blockIdx.x = (unsigned)(b % grid.x);
        blockIdx.y = (unsigned)(b / grid.x % grid.y);
        blockIdx.z = (unsigned)(b / grid.x / grid.y);
        if (!sync) {
            for (unsigned t = 0; t < block; ++t) {
                threadIdx.x = t;
                body();
            }
            continue;
        }
        for (unsigned t = 0; t < block; ++t) fibers[t].st = RUN;
        do {
            for (unsigned t = 0; t < block; ++t) {
                if (fibers[t].st != RUN) continue;
                cur = (int)t;
                threadIdx.x = t;
                if (!_setjmp(sched_jb)) {
                    Fiber &f = fibers[t];
                    if (!f.started) {
                        if (!f.stack) f.stack = (char *)malloc(STACK);
                        getcontext(&f.ctx);
                        f.ctx.uc_stack.ss_sp = f.stack;
                        f.ctx.uc_stack.ss_size = STACK;
                        f.ctx.uc_link = nullptr;
                        makecontext(&f.ctx, trampoline, 0);
                        f.started = true;
                        swapcontext(&sched_ctx, &f.ctx);
                    } else {
                        _longjmp(f.jb, 1);
                    }
                }
            }
        } while (release(block));
    }
    fiber_mode = false;
    lds_base = nullptr;
}

}  // namespace emu
