// CPU SIMT runtime behind tests/simt/include/hip/hip_runtime.h: fibers per GPU
// thread, wave intrinsics as meeting points, one block at a time.  Test
// infrastructure only.
#include <hip/hip_runtime.h>

#include <stdlib.h>
#include <sys/mman.h>

#include <chrono>
#include <vector>

#if defined(__SANITIZE_ADDRESS__)
#include <sanitizer/common_interface_defs.h>
#define SIM_ASAN 1
#else
#define SIM_ASAN 0
#endif

extern "C" void sim_switch(void **save_sp, void *load_sp);
asm(R"(
.text
.globl sim_switch
.type sim_switch,@function
sim_switch:
  pushq %rbp
  pushq %rbx
  pushq %r12
  pushq %r13
  pushq %r14
  pushq %r15
  movq %rsp, (%rdi)
  movq %rsi, %rsp
  popq %r15
  popq %r14
  popq %r13
  popq %r12
  popq %rbx
  popq %rbp
  ret
.size sim_switch, .-sim_switch
)");

namespace sim {

enum State { RUN = 0, WAIT_WAVE = 1, WAIT_BAR = 2, DONE = 3 };

struct Lane {
  void *sp = nullptr;
  char *stack = nullptr;
  size_t stack_size = 0;
  unsigned tid = 0, lane = 0, wave = 0;
  int state = DONE;
  int site = 0, op = 0, arg = 0, width = 64;
  uint64_t val = 0, result = 0;
};

Lane *g_cur = nullptr;
dim3 g_block_idx, g_block_dim, g_grid_dim;

namespace {
const size_t kStack = 256 * 1024;
std::vector<Lane> g_lanes;
void *g_sched_sp = nullptr;
const void *g_sched_lo = nullptr;
size_t g_sched_size = 0;
const std::function<void()> *g_body = nullptr;
long long g_diag_inactive = 0;  // reads of lanes outside the active set
// waves run first to last between meeting points, or last to first
// (SIMT_ORDER=reverse): either order is a schedule the GPU may take, and
// the two expose the two directions of an unsynchronised write/read between
// waves
int g_reverse = -1;

void to_sched(Lane *l, bool dying) {
#if SIM_ASAN
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(dying ? nullptr : &fake, g_sched_lo, g_sched_size);
  sim_switch(&l->sp, g_sched_sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#else
  (void)dying;
  sim_switch(&l->sp, g_sched_sp);
#endif
}

void fiber_entry() {
#if SIM_ASAN
  __sanitizer_finish_switch_fiber(nullptr, &g_sched_lo, &g_sched_size);
#endif
  Lane *l = g_cur;
  (*g_body)();
  l->state = DONE;
  to_sched(l, true);
  abort();  // never resumed
}

void resume(Lane *l) {
  g_cur = l;
  l->state = RUN;
#if SIM_ASAN
  void *fake = nullptr;
  __sanitizer_start_switch_fiber(&fake, l->stack, l->stack_size);
  sim_switch(&g_sched_sp, l->sp);
  __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#else
  sim_switch(&g_sched_sp, l->sp);
#endif
  g_cur = nullptr;
}

void start(Lane *l) {
  // a fresh stack whose first switch "returns" into fiber_entry
  uintptr_t top = ((uintptr_t)(l->stack + l->stack_size)) & ~(uintptr_t)15;
  void **sp = (void **)top;
  *--sp = nullptr;                      // fake return address of fiber_entry
  *--sp = (void *)&fiber_entry;         // popped by sim_switch's ret
  for (int k = 0; k < 6; k++) *--sp = nullptr;  // rbp rbx r12 r13 r14 r15
  l->sp = sp;
  resume(l);
}

void park(int state) {
  Lane *l = g_cur;
  l->state = state;
  to_sched(l, false);
}

// the value a lane of the active set reads from lane `src`
inline uint64_t from_lane(Lane *const *wave, const char *act, int src) {
  if (src < 0 || src >= 64 || !wave[src] || !act[src]) {
    g_diag_inactive++;
    return 0;
  }
  return wave[src]->val;
}

void release_group(Lane *const *wave, int site, int op) {
  char act[64];
  uint64_t ballot = 0;
  int first = -1;
  for (int q = 0; q < 64; q++) {
    Lane *l = wave[q];
    act[q] = l && l->state == WAIT_WAVE && l->site == site && l->op == op;
    if (act[q]) {
      if (first < 0) first = q;
      if (l->val) ballot |= 1ull << q;
    }
  }
  for (int q = 0; q < 64; q++) {
    if (!act[q]) continue;
    Lane *l = wave[q];
    const int self = q, w = l->width;
    switch (op) {
      case OP_BALLOT: l->result = ballot; break;
      case OP_SHFL: {
        int idx = (l->arg & (w - 1)) + (self & ~(w - 1));
        l->result = from_lane(wave, act, idx & 63);
        break;
      }
      case OP_SHFL_UP: {
        int idx = self - l->arg;
        if (idx < (self & ~(w - 1))) idx = self;
        l->result = from_lane(wave, act, idx);
        break;
      }
      case OP_SHFL_XOR: {
        int idx = self ^ l->arg;
        if (idx >= ((self + w) & ~(w - 1))) idx = self;
        l->result = from_lane(wave, act, idx);
        break;
      }
      case OP_READLANE: l->result = from_lane(wave, act, l->arg & 63); break;
      case OP_READFIRST: l->result = wave[first]->val; break;
      default: l->result = 0; break;
    }
  }
  for (int q = 0; q < 64; q++)
    if (act[q]) resume(wave[q]);
}
}  // namespace

unsigned lane_id() { return g_cur ? g_cur->lane : 0; }
unsigned lane_tid(const Lane *l) { return l->tid; }

uint64_t wave_op(int op, int site, uint64_t v, int arg, int width) {
  Lane *l = g_cur;
  l->op = op;
  l->site = site;
  l->val = v;
  l->arg = arg;
  l->width = width;
  park(WAIT_WAVE);
  return l->result;
}

int block_barrier(int site, int pred) {
  Lane *l = g_cur;
  l->site = site;
  l->val = pred ? 1 : 0;
  park(WAIT_BAR);
  return (int)l->result;
}

static void run_block(unsigned nthreads) {
  if (g_lanes.size() < nthreads) {
    size_t old = g_lanes.size();
    g_lanes.resize(nthreads);
    for (size_t q = old; q < nthreads; q++) {
      g_lanes[q].stack = (char *)malloc(kStack);
      g_lanes[q].stack_size = kStack;
    }
  }
  if (g_reverse < 0) {
    const char *e = getenv("SIMT_ORDER");
    g_reverse = e && !strcmp(e, "reverse");
  }
  const unsigned nw = (nthreads + 63) / 64;
  static Lane *waves[16][64];
  for (unsigned t = 0; t < nthreads; t++) {
    Lane &l = g_lanes[t];
    l.tid = t;
    l.lane = t % 64;
    l.wave = t / 64;
    l.state = RUN;
    waves[l.wave][l.lane] = &l;
  }
  for (unsigned q = 0; q < nw; q++) {
    const unsigned w = g_reverse ? nw - 1 - q : q;
    for (unsigned t = w * 64; t < nthreads && t < w * 64 + 64; t++) start(&g_lanes[t]);
  }
  for (;;) {
    bool progress = false;
    for (unsigned q = 0; q < nw; q++) {
      const unsigned w = g_reverse ? nw - 1 - q : q;
      for (;;) {
        int site = 0x7fffffff, op = -1;
        for (int q2 = 0; q2 < 64; q2++) {
          Lane *l = waves[w][q2];
          if (l->state == WAIT_WAVE && l->site < site) {
            site = l->site;
            op = l->op;
          }
        }
        if (op < 0) break;
        release_group(waves[w], site, op);
        progress = true;
      }
    }
    if (progress) continue;
    // no wave waits at an intrinsic: a block barrier, or the end
    bool any_bar = false, all = true;
    int orv = 0;
    for (unsigned t = 0; t < nthreads; t++) {
      Lane &l = g_lanes[t];
      if (l.state == WAIT_BAR) {
        any_bar = true;
        orv |= (int)l.val;
      } else if (l.state != DONE) {
        all = false;
      }
    }
    if (!any_bar) break;
    if (!all) {
      fprintf(stderr, "simt: barrier deadlock\n");
      abort();
    }
    for (unsigned q = 0; q < nthreads; q++) {
      // waves in the chosen order, the lanes of a wave in lane order
      const unsigned t = g_reverse ? (nw - 1 - q / 64) * 64 + q % 64 : q;
      if (t >= nthreads) continue;
      Lane &l = g_lanes[t];
      if (l.state == WAIT_BAR) {
        l.result = (uint64_t)orv;
        resume(&l);
      }
    }
  }
}

// graph capture: launches on a capturing stream are recorded, not run
struct Rec {
  dim3 g, b;
  std::function<void()> body;
};
}  // namespace sim

struct sim_stream {
  int capturing = 0;
  std::vector<sim::Rec> rec;
};
struct sim_graph {
  std::vector<sim::Rec> rec;
};
struct sim_event {
  std::chrono::steady_clock::time_point t;
};

namespace sim {
static void run_launch(dim3 grid, dim3 block, const std::function<void()> &body) {
  const unsigned nthreads = block.x * block.y * block.z;
  if (block.y != 1 || block.z != 1 || nthreads == 0 || nthreads > 1024 || nthreads % 64) {
    fprintf(stderr, "simt: unsupported block shape %u x %u x %u\n", block.x, block.y, block.z);
    abort();
  }
  g_grid_dim = grid;
  g_block_dim = block;
  g_body = &body;
  for (unsigned z = 0; z < grid.z; z++)
    for (unsigned y = 0; y < grid.y; y++)
      for (unsigned x = 0; x < grid.x; x++) {
        g_block_idx = dim3(x, y, z);
        run_block(nthreads);
      }
  g_body = nullptr;
}

void launch(dim3 grid, dim3 block, void *stream, std::function<void()> body) {
  sim_stream *st = (sim_stream *)stream;
  if (st && st->capturing) {
    st->rec.push_back(Rec{grid, block, std::move(body)});
    return;
  }
  run_launch(grid, block, body);
}
}  // namespace sim

sim_tidx sim_thread_idx() { return sim_tidx{sim::g_cur ? sim::g_cur->tid : 0u, 0u, 0u}; }

extern "C" long long simt_inactive_reads() { return sim::g_diag_inactive; }

hipError_t hipMalloc(void **p, size_t n) {
  *p = malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
  free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
  if (n) memmove(d, s, n);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st) {
  if (st && st->capturing) {
    fprintf(stderr, "simt: copy recorded into a graph\n");
    abort();
  }
  return hipMemcpy(d, s, n, k);
}
hipError_t hipMemcpy2DAsync(void *d, size_t dp, const void *s, size_t sp, size_t w, size_t h,
                            hipMemcpyKind, hipStream_t) {
  for (size_t r = 0; r < h; r++) memmove((char *)d + r * dp, (const char *)s + r * sp, w);
  return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n) {
  if (n) memset(d, v, n);
  return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st) {
  if (st && st->capturing) {
    fprintf(stderr, "simt: memset recorded into a graph\n");
    abort();
  }
  return hipMemset(d, v, n);
}
hipError_t hipStreamCreateWithFlags(hipStream_t *st, unsigned) {
  *st = new sim_stream();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t st) {
  delete st;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
void sincospi(double x, double *s, double *c) {
  *s = sin(M_PI * x);
  *c = cos(M_PI * x);
}
hipError_t hipGetDevice(int *dev) {
  *dev = 0;
  return hipSuccess;
}
hipError_t hipGetSymbolAddress(void **p, const void *sym) {
  *p = (void *)sym;
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t st, int) {
  st->capturing = 1;
  st->rec.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t st, hipGraph_t *g) {
  *g = new sim_graph();
  (*g)->rec = std::move(st->rec);
  st->rec.clear();
  st->capturing = 0;
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  *e = new sim_graph();
  (*e)->rec = g->rec;
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (auto &r : e->rec) sim::run_launch(r.g, r.b, r.body);
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new sim_event();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
const char *hipGetErrorString(hipError_t) { return "simt error"; }
