// ipg_ensemble.hip -- a batch of independent cases advanced through each
// substep together (ipg_batch_* in include/ipg.h).
//
// A batch records one substep of every case: each launch the case's own
// substep would issue becomes a record (its body, grid, block, dynamic LDS and
// argument struct) tagged with its substep stage.  Records of one stage are
// merged across cases into one launch per kernel (per FFT length, whose plans
// are compile-time): k_batch<K> over a merged index space in which each case
// owns a run of blocks and reads its own argument struct (ipg_batch.h).  The
// merged launches of a substep -- about 20 for any number of cases, against
// about 20 per case -- are captured as one hipGraph per (RK stage, buffer
// parity).  Records are made once per key and kept on the device; a case's
// configuration counter (ipg_ctx::gen) moving makes the batch record again.
// The cases may be split into groups on streams of their own, whose merged
// launches run side by side inside the one graph.
#include "ipg_internal.h"
#include <algorithm>
#include <map>
#include <string>
#include <vector>

struct IpgRecEntry {
  BatchFns fns;
  dim3 grid, block;
  size_t lds;
  std::vector<unsigned char> args;
  std::string stage;
  bool side;  // issued on the side stream (the carried Newton-Euler update)
  bool join;  // not a launch: the main stream waits for the side stream here
};
struct IpgRec {
  std::vector<IpgRecEntry> e;
};

int ipg_rec_add(IpgRec *r, BatchFns fns, dim3 grid, dim3 block, size_t lds, const void *args, size_t argsz,
                const char *stage, bool side) {
  const unsigned char *a = (const unsigned char *)args;
  r->e.push_back(IpgRecEntry{fns, grid, block, lds, std::vector<unsigned char>(a, a + argsz), stage ? stage : "?",
                             side, false});
  return 0;
}

void ipg_rec_join(IpgRec *r) {
  r->e.push_back(IpgRecEntry{BatchFns{nullptr, nullptr}, dim3(1), dim3(1), 0, {}, "join", false, true});
}

namespace {

// (rkiter - 1) + 3 parity + 6 carried: the substep of RK stage rkiter from
// c/dcdt buffer parity `parity`, carrying the previous substep's Newton-Euler
// update beside its RK stage or not
constexpr int NKEY = 12;

// One merged launch: the cases holding a record of one stage with one body
// (or, with join set, the point where the main stream waits for the side).
struct Merged {
  BatchFns fns;
  bool side, join;
  dim3 block;
  size_t lds;
  int total;            // merged blocks (a multiple of 8)
  int ncase;
  size_t off_args, off_seg, off_of;  // in the plan's device table
  std::string stage;    // stage#occurrence
  std::string tune_key; // stage#occurrence|body: the tuner's choice is kept per key
};

struct Plan {
  std::vector<Merged> L;
  void *dev = nullptr;  // argument structs, segments and block->segment tables of every launch
};

std::string hexptr(const void *p) {
  char b[32];
  std::snprintf(b, sizeof b, "%p", p);
  return b;
}

}  // namespace

// The cases are split into groups, each with its own stream and merged
// launches: the groups' substeps run side by side, so that the latency-bound
// stages of one group (the IBM passes, the Thomas solve, the sphere integrals)
// overlap the bandwidth-bound stages of another.  One group is one stream of
// merged launches; the tuner chooses the number of groups.
struct Group {
  std::vector<int> idx;          // the cases it holds, in batch order
  hipStream_t stream = nullptr;  // the batch's stream for group 0
  hipStream_t side = nullptr;    // the carried updates beside the RK stage
  hipEvent_t joined = nullptr;   // its substep done (another group's stream)
  hipEvent_t fork = nullptr, side_done = nullptr;
  Plan plan[NKEY];
};

struct ipg_batch {
  std::vector<ipg_ctx *> cs;
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr, fork = nullptr;
  std::vector<hipEvent_t> case_ev;
  std::vector<unsigned long> seen;  // each case's entry-point count when the batch last waited for its stream
  std::vector<hipEvent_t> tev;  // timing events
  bool use_graphs = true;
  int ncu = 256;
  std::vector<Group> groups;
  hipGraphExec_t graph[NKEY] = {};
  std::map<std::string, int> persist;  // tune_key -> persistent workgroups per CU (0: one per block)
  int persist_default = 0;             // for a launch the tuner has not seen
  std::map<const void *, size_t> lds_set;  // the dynamic-LDS limit each merged kernel has been given
  std::vector<unsigned> gen;           // the cases' configuration counters the plans were recorded at
  // defer each case's Newton-Euler update to the next substep, which carries
  // it on a side stream beside its RK stage (the tuner's choice); off, the
  // update closes its substep
  bool carry = true;
  bool carry_planned = true;           // the carry mode the plans were recorded in
};

namespace {

void drop_plan(Plan &P) {
  if (P.dev) (void)hipFree(P.dev);
  P = Plan{};
}

void drop_graphs(ipg_batch *b) {
  for (hipGraphExec_t &g : b->graph)
    if (g) { (void)hipGraphExecDestroy(g); g = nullptr; }
}

void drop_groups(ipg_batch *b) {
  drop_graphs(b);
  for (Group &G : b->groups) {
    for (Plan &P : G.plan) drop_plan(P);
    if (G.stream && G.stream != b->stream) (void)hipStreamDestroy(G.stream);
    if (G.side) (void)hipStreamDestroy(G.side);
    for (hipEvent_t e : {G.joined, G.fork, G.side_done})
      if (e) (void)hipEventDestroy(e);
  }
  b->groups.clear();
}

// ng groups of about equal cell counts: cases by decreasing size, each to the
// group holding the fewest cells so far
int make_groups(ipg_batch *b, int ng) {
  IPG_HIP(hipDeviceSynchronize());
  drop_groups(b);
  const int n = (int)b->cs.size();
  ng = std::max(1, std::min(ng, n));
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  auto cells = [&](int i) { const Grid &g = b->cs[i]->g; return (double)g.nx * g.ny * g.nz; };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cells(x) > cells(y); });
  std::vector<double> load(ng, 0.0);
  b->groups.resize(ng);
  for (int i : order) {
    const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b->groups[g].idx.push_back(i);
    load[g] += cells(i);
  }
  for (int g = 0; g < ng; g++) {
    Group &G = b->groups[g];
    std::sort(G.idx.begin(), G.idx.end());
    if (g == 0) G.stream = b->stream;
    else IPG_HIP(hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking));
    IPG_HIP(hipStreamCreateWithFlags(&G.side, hipStreamNonBlocking));
    IPG_HIP(hipEventCreateWithFlags(&G.joined, hipEventDisableTiming));
    IPG_HIP(hipEventCreateWithFlags(&G.fork, hipEventDisableTiming));
    IPG_HIP(hipEventCreateWithFlags(&G.side_done, hipEventDisableTiming));
  }
  return 0;
}

// true when no case has changed its configuration since the plans were recorded
bool gen_current(const ipg_batch *b) {
  if (b->gen.size() != b->cs.size()) return false;
  for (size_t i = 0; i < b->cs.size(); i++)
    if (b->gen[i] != b->cs[i]->gen) return false;
  return true;
}

// Records one substep of every case of group G for key r and merges the
// records.
int build_plan(ipg_batch *b, Group &G, int r) {
  const int rkiter = r % 3 + 1, par = (r / 3) % 2;
  const bool carried = r >= 6;
  const int n = (int)G.idx.size();
  drop_plan(G.plan[r]);
  std::vector<IpgRec> rec(n);
  for (int i = 0; i < n; i++)
    if (ipg_batch_record(b->cs[G.idx[i]], rkiter, par, carried, b->carry, &rec[i])) return 1;
  // stage#occurrence keys per case; the longest sequence gives the order, and
  // every other case's sequence must be a subsequence of it (a case without
  // spheres has no IBM or particle stages)
  std::vector<std::map<std::string, int>> at(n);
  std::vector<std::vector<std::string>> keys(n);
  int ref = 0;
  for (int i = 0; i < n; i++) {
    std::map<std::string, int> occ;
    for (size_t j = 0; j < rec[i].e.size(); j++) {
      const std::string k = rec[i].e[j].stage + "#" + std::to_string(occ[rec[i].e[j].stage]++);
      keys[i].push_back(k);
      at[i][k] = (int)j;
    }
    if (keys[i].size() > keys[ref].size()) ref = i;
  }
  std::map<std::string, int> pos;
  for (size_t j = 0; j < keys[ref].size(); j++) pos[keys[ref][j]] = (int)j;
  for (int i = 0; i < n; i++) {
    int last = -1;
    for (const std::string &k : keys[i]) {
      auto it = pos.find(k);
      if (it == pos.end() || it->second <= last) {
        ipg_set_error("batch: the cases' substeps do not run the same stages in the same order (" + k + ")");
        return 1;
      }
      last = it->second;
    }
  }
  Plan &P = G.plan[r];
  std::vector<unsigned char> blob;
  auto align = [&](size_t a) { blob.resize((blob.size() + a - 1) / a * a); };
  auto put = [&](const void *p, size_t nb) {
    const unsigned char *q = (const unsigned char *)p;
    blob.insert(blob.end(), q, q + nb);
  };
  for (const std::string &k : keys[ref]) {
    // the cases holding stage k, grouped by body (kernel), in case order
    std::vector<std::pair<const void *, std::vector<int>>> groups;
    for (int i = 0; i < n; i++) {
      auto it = at[i].find(k);
      if (it == at[i].end()) continue;
      const void *fn = rec[i].e[it->second].fns.one;
      auto g = std::find_if(groups.begin(), groups.end(), [&](const auto &x) { return x.first == fn; });
      if (g == groups.end()) groups.push_back({fn, {i}});
      else g->second.push_back(i);
    }
    for (const auto &grp : groups) {
      const IpgRecEntry &e0 = rec[grp.second[0]].e[at[grp.second[0]][k]];
      Merged M{};
      M.fns = e0.fns;
      M.side = e0.side;
      M.join = e0.join;
      M.stage = k;
      for (int i : grp.second)
        if (rec[i].e[at[i][k]].side != e0.side) {
          ipg_set_error("batch: the cases run a stage on different streams (" + k + ")");
          return 1;
        }
      if (M.join) { P.L.push_back(M); continue; }
      M.block = e0.block;
      M.lds = 0;
      M.ncase = (int)grp.second.size();
      M.stage = k;
      M.tune_key = k + "|" + hexptr(e0.fns.one);
      const size_t argsz = e0.args.size();
      align(16);
      M.off_args = blob.size();
      for (int i : grp.second) {
        const IpgRecEntry &e = rec[i].e[at[i][k]];
        if (e.args.size() != argsz || e.block.x != M.block.x || e.block.y != M.block.y || e.block.z != M.block.z) {
          ipg_set_error("batch: records of one kernel differ in their block or argument size");
          return 1;
        }
        put(e.args.data(), argsz);
        M.lds = std::max(M.lds, e.lds);
      }
      align(16);
      M.off_seg = blob.size();
      long running = 0;
      std::vector<int> of;
      for (int s = 0; s < M.ncase; s++) {
        const IpgRecEntry &e = rec[grp.second[s]].e[at[grp.second[s]][k]];
        const BatchSeg sg{(int)running, (int)e.grid.x, (int)e.grid.y, (int)e.grid.z};
        put(&sg, sizeof sg);
        const long cnt = (long)e.grid.x * e.grid.y * e.grid.z;
        const long run8 = (cnt + 7) / 8 * 8;
        for (long t = 0; t < run8 / 8; t++) of.push_back(s);
        running += run8;
        if (running >= (1L << 30)) { ipg_set_error("batch: a merged launch of more than 2^30 blocks"); return 1; }
      }
      M.total = (int)running;
      align(16);
      M.off_of = blob.size();
      put(of.data(), of.size() * sizeof(int));
      if (M.lds > 0)
        for (const void *fn : {M.fns.one, M.fns.persist})
          if (b->lds_set[fn] < M.lds) {
            IPG_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)M.lds));
            b->lds_set[fn] = M.lds;
          }
      P.L.push_back(M);
    }
  }
  IPG_HIP(hipMalloc(&P.dev, std::max<size_t>(blob.size(), 16)));
  IPG_HIP(hipMemcpy(P.dev, blob.data(), blob.size(), hipMemcpyHostToDevice));
  return 0;
}

// The merged launches of plan P for group G: the main launches on stream s,
// the side launches forked onto G.side and joined back where the plan says.
// With timing, every launch goes on s in the plan's order with an event after
// each (b->tev[t0 + q + 1], b->tev[t0] before the first; joins take none).
int launch_merged(ipg_batch *b, const Plan &P, const Merged &M, hipStream_t s) {
  const char *base = (const char *)P.dev;
  const void *args = base + M.off_args;
  BatchMap m{(const BatchSeg *)(base + M.off_seg), (const int *)(base + M.off_of), M.total};
  auto it = b->persist.find(M.tune_key);
  const int f = it == b->persist.end() ? b->persist_default : it->second;
  unsigned grid = (unsigned)M.total;
  const void *fn = M.fns.one;
  if (f > 0) {
    const long cap = std::max<long>(8, (long)f * b->ncu / 8 * 8);
    if (cap < M.total) { grid = (unsigned)cap; fn = M.fns.persist; }
  }
  void *kp[] = {(void *)&args, (void *)&m};
  IPG_HIP(hipLaunchKernel(fn, dim3(grid), M.block, kp, M.lds, s));
  return 0;
}

int run_plan(ipg_batch *b, Group &G, const Plan &P, hipStream_t s, bool timing, size_t t0 = 0) {
  if (timing) {
    while (b->tev.size() < t0 + P.L.size() + 1) {
      hipEvent_t e;
      IPG_HIP(hipEventCreate(&e));
      b->tev.push_back(e);
    }
    IPG_HIP(hipEventRecord(b->tev[t0], s));
    size_t q = 0;
    for (const Merged &M : P.L) {
      if (M.join) continue;
      if (launch_merged(b, P, M, s)) return 1;
      IPG_HIP(hipEventRecord(b->tev[t0 + ++q], s));
    }
    return 0;
  }
  bool forked = false;
  for (const Merged &M : P.L) {
    if (M.join) {
      if (forked) {
        IPG_HIP(hipEventRecord(G.side_done, G.side));
        IPG_HIP(hipStreamWaitEvent(s, G.side_done, 0));
        forked = false;
      }
      continue;
    }
    if (M.side && !forked) {
      IPG_HIP(hipEventRecord(G.fork, s));
      IPG_HIP(hipStreamWaitEvent(G.side, G.fork, 0));
      forked = true;
    }
    if (launch_merged(b, P, M, M.side ? G.side : s)) return 1;
  }
  if (forked) {
    IPG_HIP(hipEventRecord(G.side_done, G.side));
    IPG_HIP(hipStreamWaitEvent(s, G.side_done, 0));
  }
  return 0;
}

int ensure_plans(ipg_batch *b, int r) {
  if (b->groups.empty() && make_groups(b, 1)) return 1;
  if (!gen_current(b) || b->carry_planned != b->carry) {
    // a case changed, or the carry mode: every key records again
    b->carry_planned = b->carry;
    IPG_HIP(hipDeviceSynchronize());
    drop_graphs(b);
    for (Group &G : b->groups)
      for (Plan &P : G.plan) drop_plan(P);
    b->gen.clear();
    for (ipg_ctx *c : b->cs) b->gen.push_back(c->gen);
  }
  for (Group &G : b->groups)
    if (!G.plan[r].dev) {
      IPG_HIP(hipDeviceSynchronize());
      if (build_plan(b, G, r)) return 1;
    }
  return 0;
}

// The groups' merged launches of key r: group 0 on the batch's stream, the
// others forked from it and joined back.
int run_groups(ipg_batch *b, int r) {
  if (b->groups.size() > 1) IPG_HIP(hipEventRecord(b->fork, b->stream));
  for (size_t g = 1; g < b->groups.size(); g++) IPG_HIP(hipStreamWaitEvent(b->groups[g].stream, b->fork, 0));
  for (Group &G : b->groups)
    if (run_plan(b, G, G.plan[r], G.stream, false)) return 1;
  for (size_t g = 1; g < b->groups.size(); g++) {
    IPG_HIP(hipEventRecord(b->groups[g].joined, b->groups[g].stream));
    IPG_HIP(hipStreamWaitEvent(b->stream, b->groups[g].joined, 0));
  }
  return 0;
}

int capture(ipg_batch *b, int r) {
  hipGraph_t graph;
  IPG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = run_groups(b, r);
  const hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (rc) return rc;
  if (e != hipSuccess) { ipg_set_error(std::string("batch graph capture: ") + hipGetErrorString(e)); return 1; }
  IPG_HIP(hipGraphInstantiate(&b->graph[r], graph, nullptr, nullptr, 0));
  (void)hipGraphDestroy(graph);
  return 0;
}

// A merged launch of a timed substep and the event recorded before it (the
// next event follows it).
struct Timed {
  const Merged *m;
  size_t ev;
};

// One substep of every case.  With timed: eager, the groups one after
// another on the batch's stream, an event after every merged launch, and
// *timed gets the launches in that order.
int substep(ipg_batch *b, int rkiter, std::vector<Timed> *timed) {
  if (rkiter < 1 || rkiter > 3) { ipg_set_error("batch: rkiter must be 1..3"); return 1; }
  const int n = (int)b->cs.size();
  const int par = b->cs[0]->cpar;
  for (int i = 1; i < n; i++)
    if (b->cs[i]->cpar != par) {
      ipg_set_error("batch: the cases are at different c/dcdt buffer parities (an odd difference in substeps run)");
      return 1;
    }
  // the previous merged substep's Newton-Euler updates ride along beside this
  // RK stage when every case with spheres still has its RK3 predecessor's
  // pending; otherwise each pending one runs first, on its case's stream
  bool carried = false, all = true;
  for (ipg_ctx *c : b->cs) {
    all = all && ipg_batch_carries(c, rkiter);
    carried = carried || c->pc.np > 0;
  }
  carried = carried && all && b->carry;
  for (int i = 0; i < n; i++)
    if (ipg_batch_enter(b->cs[i], carried)) return 1;
  const int r = (rkiter - 1) + 3 * par + (carried ? 6 : 0);
  if (ensure_plans(b, r)) return 1;
  // the batch's work on a case follows the case's own, where there was any
  // since the batch last waited for it
  for (int i = 0; i < n; i++) {
    ipg_ctx *c = b->cs[i];
    if (c->calls == b->seen[i]) continue;
    IPG_HIP(hipEventRecord(b->case_ev[i], c->stream));
    IPG_HIP(hipStreamWaitEvent(b->stream, b->case_ev[i], 0));
    b->seen[i] = c->calls;
  }
  if (timed) {
    size_t t0 = 0;
    timed->clear();
    for (Group &G : b->groups) {
      if (run_plan(b, G, G.plan[r], b->stream, true, t0)) return 1;
      size_t q = 0;
      for (const Merged &M : G.plan[r].L)
        if (!M.join) timed->push_back(Timed{&M, t0 + q++});
      t0 += q + 1;
    }
  } else if (b->use_graphs) {
    if (!b->graph[r] && capture(b, r)) return 1;
    IPG_HIP(hipGraphLaunch(b->graph[r], b->stream));
  } else if (run_groups(b, r)) {
    return 1;
  }
  // and the case's later work follows the batch's: the next entry point on
  // the case makes its stream wait for this event
  IPG_HIP(hipEventRecord(b->done, b->stream));
  for (int i = 0; i < n; i++) ipg_batch_leave(b->cs[i], rkiter, b->carry, b->done);
  return 0;
}

// Per-launch times (ms) of a timed substep.
int launch_times(ipg_batch *b, const std::vector<Timed> &timed, std::vector<float> &t) {
  IPG_HIP(hipStreamSynchronize(b->stream));
  t.assign(timed.size(), 0.f);
  for (size_t q = 0; q < timed.size(); q++)
    IPG_HIP(hipEventElapsedTime(&t[q], b->tev[timed[q].ev], b->tev[timed[q].ev + 1]));
  return 0;
}

// ms per substep of one RK3 step replayed from the graphs (after one
// untimed step, which captures them)
int time_step(ipg_batch *b, float &ms, int &count) {
  for (int rk = 1; rk <= 3; rk++, count++)
    if (substep(b, rk, nullptr)) return 1;
  if (b->tev.size() < 2) {
    hipEvent_t e[2];
    IPG_HIP(hipEventCreate(&e[0]));
    IPG_HIP(hipEventCreate(&e[1]));
    b->tev.insert(b->tev.end(), e, e + 2);
  }
  IPG_HIP(hipEventRecord(b->tev[0], b->stream));
  for (int rk = 1; rk <= 3; rk++, count++)
    if (substep(b, rk, nullptr)) return 1;
  IPG_HIP(hipEventRecord(b->tev[1], b->stream));
  IPG_HIP(hipEventSynchronize(b->tev[1]));
  IPG_HIP(hipEventElapsedTime(&ms, b->tev[0], b->tev[1]));
  ms /= 3.f;
  return 0;
}

}  // namespace

extern "C" {

int ipg_batch_create(ipg_ctx *const *cases, int ncases, ipg_batch **out) {
  if (!cases || ncases < 1 || !out) { ipg_set_error("batch_create: bad arguments"); return 1; }
  *out = nullptr;
  for (int i = 0; i < ncases; i++)
    for (int j = 0; j < i; j++)
      if (!cases[i] || cases[i] == cases[j]) { ipg_set_error("batch_create: a null or repeated case"); return 1; }
  ipg_batch *b = new ipg_batch();
  b->cs.assign(cases, cases + ncases);
  auto fail = [&]() { ipg_batch_destroy(b); return 1; };
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&b->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    ipg_set_error("batch_create: device query failed");
    return fail();
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
    ipg_set_error("batch_create: stream or event create failed");
    return fail();
  }
  b->case_ev.assign(ncases, nullptr);
  // (unsigned long)-1: wait for every case's stream before the first substep
  b->seen.assign(ncases, (unsigned long)-1);
  for (int i = 0; i < ncases; i++)
    if (hipEventCreateWithFlags(&b->case_ev[i], hipEventDisableTiming) != hipSuccess) {
      ipg_set_error("batch_create: event create failed");
      return fail();
    }
  if (make_groups(b, 1)) return fail();
  *out = b;
  return 0;
}

int ipg_batch_destroy(ipg_batch *b) {
  if (!b) return 0;
  if (b->stream) (void)hipDeviceSynchronize();
  // the cases no longer wait for the batch's event (its work is done)
  for (ipg_ctx *c : b->cs)
    if (c->batch_pending && c->batch_ev == b->done) c->batch_pending = false;
  drop_groups(b);
  for (hipEvent_t e : b->case_ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->tev) (void)hipEventDestroy(e);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return 0;
}

int ipg_batch_substep(ipg_batch *b, int rkiter) {
  if (!b) { ipg_set_error("batch: null batch"); return 1; }
  return substep(b, rkiter, nullptr);
}

int ipg_batch_sync(ipg_batch *b) {
  if (!b) { ipg_set_error("batch: null batch"); return 1; }
  IPG_HIP(hipStreamSynchronize(b->stream));
  for (ipg_ctx *c : b->cs)
    if (ipg_sync(c)) return 3;
  return 0;
}

int ipg_batch_use_graphs(ipg_batch *b, int on) {
  if (!b) { ipg_set_error("batch: null batch"); return 1; }
  b->use_graphs = on != 0;
  return 0;
}

int ipg_batch_set_groups(ipg_batch *b, int ngroups) {
  if (!b || ngroups < 1) { ipg_set_error("batch_set_groups: bad arguments"); return 1; }
  if ((int)b->groups.size() == std::min(ngroups, (int)b->cs.size())) return 0;
  return make_groups(b, ngroups);
}

int ipg_batch_autotune(ipg_batch *b, int *substeps) {
  if (!b) { ipg_set_error("batch: null batch"); return 1; }
  int count = 0;
  std::vector<float> t;
  std::vector<Timed> timed;
  const int ng0 = (int)b->groups.size();
  if (ng0 != 1 && make_groups(b, 1)) return 1;
  // the RK stage's z planes per workgroup, the same for every case, timed
  // over one RK3 step each
  int best_kb = 0;
  float best_ms = 0.f;
  for (int kb = 16; kb <= 64; kb *= 2) {
    for (ipg_ctx *c : b->cs) {
      const int k = std::min(kb, c->g.nz);
      if (c->tune_rk_kb != k) { c->tune_rk_kb = k; c->gen++; }
    }
    float ms = 0.f;
    for (int rk = 1; rk <= 3; rk++, count++) {
      if (substep(b, rk, &timed) || launch_times(b, timed, t)) return 1;
      for (size_t q = 0; q < timed.size(); q++)
        if (timed[q].m->stage.rfind("rk_stage#", 0) == 0) ms += t[q];
    }
    if (best_kb == 0 || ms < best_ms) { best_kb = kb; best_ms = ms; }
  }
  for (ipg_ctx *c : b->cs) {
    const int k = std::min(best_kb, c->g.nz);
    if (c->tune_rk_kb != k) { c->tune_rk_kb = k; c->gen++; }
  }
  // per merged launch: one workgroup per block, or a persistent grid of f
  // workgroups per CU
  std::map<std::string, std::pair<int, float>> best;  // tune_key -> (f, ms)
  b->persist.clear();
  for (int f : {0, 1, 2, 4}) {
    b->persist_default = f;
    for (int rk = 1; rk <= 3; rk++, count++) {
      if (substep(b, rk, &timed) || launch_times(b, timed, t)) return 1;
      for (size_t q = 0; q < timed.size(); q++) {
        const std::string &key = timed[q].m->tune_key;
        auto it = best.find(key);
        if (it == best.end() || t[q] < it->second.second) best[key] = {f, t[q]};
      }
    }
  }
  b->persist_default = 0;
  for (const auto &kv : best) b->persist[kv.first] = kv.second.first;
  drop_graphs(b);
  // the number of stream groups, timed on graph replays of whole steps
  int best_ng = 1;
  float best_step = 0.f;
  for (int ng = 1; ng <= std::min(4, (int)b->cs.size()); ng *= 2) {
    float ms = 0.f;
    if (make_groups(b, ng) || time_step(b, ms, count)) return 1;
    if (ng == 1 || ms < best_step) { best_ng = ng; best_step = ms; }
  }
  if ((int)b->groups.size() != best_ng && make_groups(b, best_ng)) return 1;
  // the Newton-Euler updates carried beside the next RK stage, or closing
  // their substeps
  float carry_ms[2] = {0.f, 0.f};
  for (int on = 1; on >= 0; on--) {
    b->carry = on != 0;
    if (time_step(b, carry_ms[on], count)) return 1;
  }
  b->carry = carry_ms[1] <= carry_ms[0];
  if (substeps) *substeps = count;
  return 0;
}

int ipg_batch_set_carry(ipg_batch *b, int on) {
  if (!b) { ipg_set_error("batch: null batch"); return 1; }
  b->carry = on != 0;
  return 0;
}

int ipg_batch_carry(ipg_batch *b, int *on) {
  if (!b || !on) { ipg_set_error("batch_carry: bad arguments"); return 1; }
  *on = b->carry ? 1 : 0;
  return 0;
}

int ipg_batch_launches(ipg_batch *b, int rkiter, int *nlaunch, int *npersistent) {
  if (!b || !nlaunch || rkiter < 1 || rkiter > 3) { ipg_set_error("batch_launches: bad arguments"); return 1; }
  const int par = b->cs[0]->cpar;
  // the plan of the next substep of stage rkiter: carrying updates, if built
  int r = (rkiter - 1) + 3 * par + 6;
  if (b->groups.empty() || !b->groups[0].plan[r].dev) r -= 6;
  int nl = 0, np = 0;
  for (const Group &G : b->groups)
    for (const Merged &M : G.plan[r].L) {
      if (M.join) continue;
      nl++;
      auto it = b->persist.find(M.tune_key);
      const int f = it == b->persist.end() ? b->persist_default : it->second;
      if (f > 0 && std::max<long>(8, (long)f * b->ncu / 8 * 8) < M.total) np++;
    }
  *nlaunch = nl;
  if (npersistent) *npersistent = np;
  return 0;
}

int ipg_batch_groups(ipg_batch *b, int *ngroups) {
  if (!b || !ngroups) { ipg_set_error("batch_groups: bad arguments"); return 1; }
  *ngroups = (int)b->groups.size();
  return 0;
}

int ipg_batch_phases(ipg_batch *b, int rkiter, int maxn, double *ms, const char **names, int *n) {
  if (!b || !ms || !names || !n) { ipg_set_error("batch_phases: bad arguments"); return 1; }
  std::vector<Timed> timed;
  std::vector<float> t;
  if (substep(b, rkiter, &timed) || launch_times(b, timed, t)) return 1;
  int k = 0;
  for (size_t q = 0; q < timed.size() && k < maxn; q++, k++) {
    ms[k] = t[q];
    names[k] = timed[q].m->stage.c_str();
  }
  *n = k;
  return 0;
}

}  // extern "C"
