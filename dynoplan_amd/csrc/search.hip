// search.hip -- the db-A* search loop on the batched expansion path.
//
// Host counterpart of dbastar() (src/dbastar/dbastar.cpp:523-1010): an open
// list ordered by the reference's comparator (lowest float fScore, then
// highest float gScore; dbastar.cpp:56-68), nodes owned by one vector, the
// goal test on every popped node (dbastar.cpp:787-802), expansion of the
// popped nodes by one dp_expand_batch, node creation for NOVEL pairs and
// rewiring of the neighbours of duplicate pairs (dbastar.cpp:856-912), and
// the solution trajectory (from_solution_to_yaml_and_traj,
// dbastar.cpp:129-310).  With batch_nodes = 1 the loop pops one node per
// iteration like the reference; primitives are taken in canonical order (the
// reference shuffles them).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "dp_device.h"
#include "dp_internal.h"

namespace {

struct node_rec {
  float f, g, h;
  int32_t parent;   // -1 for the start
  int32_t motion;   // used_motion
  int32_t inter;    // intermediate_state (-1: the whole primitive)
  int32_t handle;   // heap entry of the node's last push
  bool in_open;
};

// Binary max-heap of entries under compareAStarNode.  The reference re-pushes
// a closed node when it is rewired without marking it open again, so a node
// can own several entries; an entry id is a handle.
struct open_list {
  std::vector<node_rec>* nodes = nullptr;
  std::vector<int32_t> heap;       // entry ids
  std::vector<int32_t> entry_node; // entry -> node
  std::vector<int32_t> entry_pos;  // entry -> position in heap (-1 popped)

  // compareAStarNode(a, b): true when a comes after b
  bool after(int32_t ea, int32_t eb) const {
    const node_rec& a = (*nodes)[entry_node[ea]];
    const node_rec& b = (*nodes)[entry_node[eb]];
    if (a.f != b.f) return a.f > b.f;
    return a.g < b.g;
  }
  void put(size_t i, int32_t e) {
    heap[i] = e;
    entry_pos[e] = (int32_t)i;
  }
  void sift_up(size_t i) {
    while (i > 0) {
      const size_t p = (i - 1) / 2;
      if (!after(heap[p], heap[i])) break;
      const int32_t e = heap[i];
      put(i, heap[p]);
      put(p, e);
      i = p;
    }
  }
  void sift_down(size_t i) {
    const size_t n = heap.size();
    for (;;) {
      const size_t c = 2 * i + 1;
      if (c >= n) break;
      size_t top = c;  // first of the best children
      if (c + 1 < n && after(heap[c], heap[c + 1])) top = c + 1;
      if (after(heap[top], heap[i])) break;
      const int32_t e = heap[i];
      put(i, heap[top]);
      put(top, e);
      i = top;
    }
  }
  int32_t push(int32_t node) {
    const int32_t e = (int32_t)entry_node.size();
    entry_node.push_back(node);
    entry_pos.push_back((int32_t)heap.size());
    heap.push_back(e);
    sift_up(heap.size() - 1);
    return e;
  }
  int32_t pop() {
    const int32_t e = heap[0];
    put(0, heap.back());
    heap.pop_back();
    if (!heap.empty()) sift_down(0);
    entry_pos[e] = -1;
    return entry_node[e];
  }
  void increase(int32_t e) {
    if (entry_pos[e] >= 0) sift_up((size_t)entry_pos[e]);
  }
  bool empty() const { return heap.empty(); }
};

// Nodes hashed by position cell (cell edge >= r / w_position): every node
// within metric distance r of a query lies in the 3^d cells around it.
struct node_grid {
  double inv = 1.0;
  int dims = 2;
  std::unordered_map<uint64_t, std::vector<int32_t>> cells;
  static uint64_t key(int64_t a, int64_t b, int64_t c) {
    return ((uint64_t)(a & 0x1fffff) << 42) | ((uint64_t)(b & 0x1fffff) << 21) | (uint64_t)(c & 0x1fffff);
  }
  void cell_of(const double* x, int64_t* ci) const {
    for (int k = 0; k < 3; ++k) ci[k] = k < dims ? (int64_t)std::floor(x[k] * inv) : 0;
  }
  void add(const double* x, int32_t id) {
    int64_t ci[3];
    cell_of(x, ci);
    cells[key(ci[0], ci[1], ci[2])].push_back(id);
  }
  template <typename F>
  void around(const double* x, F&& f) const {
    int64_t ci[3];
    cell_of(x, ci);
    const int zr = dims > 2 ? 1 : 0;
    for (int64_t a = -1; a <= 1; ++a)
      for (int64_t b = -1; b <= 1; ++b)
        for (int64_t z = -zr; z <= zr; ++z) {
          auto it = cells.find(key(ci[0] + a, ci[1] + b, ci[2] + z));
          if (it == cells.end()) continue;
          for (int32_t id : it->second) f(id);
        }
  }
};

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// transform_primitive of motion j to the node state x (all states).
void transform(const dp_ctx* c, const double* x, int32_t j, std::vector<double>& out) {
  const int nx = c->desc.nx;
  const dp_dev_sys& S = c->hsys;
  const int32_t s0 = c->h_off[j], L1 = c->h_off[j + 1] - s0;
  out.assign((size_t)L1 * nx, 0.0);
  if (S.system != DP_SYS_QUADROTOR) {
    for (int32_t k = 0; k < L1; ++k)
      for (int d = 0; d < nx; ++d) {
        const double v = c->h_st[(size_t)(s0 + k) * nx + d];
        out[(size_t)k * nx + d] = dp_is_offset_dim(S, d) ? v + x[d] : v;
      }
    return;
  }
  for (int d = 0; d < nx; ++d) {
    const double v = c->h_st[(size_t)s0 * nx + d];
    out[d] = dp_is_offset_dim(S, d) ? v + x[d] : v;
  }
  for (int32_t k = 1; k < L1; ++k) {
    double u[4];
    for (int a = 0; a < 4; ++a) u[a] = c->h_ac[(size_t)(s0 - j + k - 1) * c->h_nu + a];
    dp_quad_step64(S, &out[(size_t)(k - 1) * nx], u, &out[(size_t)k * nx]);
  }
}

}  // namespace

extern "C" {

void dp_search_options_default(dp_search_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->delta = 0.3f;  // options.hpp:108-111
  o->alpha = 0.5f;
  o->delta_factor_goal = 1.0;
  o->cost_delta_factor = 1.0;
  o->max_expands = 1000000;
  o->search_timelimit_ms = 1e4;
  o->limit_branching_factor = 20;
  o->heuristic = 0;
  o->batch_nodes = 1;
}

int dp_search(dp_ctx* c, const double* start, const double* goal, const dp_search_options* opt,
              dp_search_info* info) {
  if (!c || !start || !goal || !opt || !info) {
    dp_set_error("null argument");
    return DP_ERR_ARG;
  }
  if (!(opt->alpha > 0.0f && opt->alpha < 1.0f)) {  // dbastar.cpp:588-590
    dp_set_error("Alpha needs to be between 0 and 1!");
    return DP_ERR_ARG;
  }
  if (opt->batch_nodes < 1 || (opt->heuristic != 0 && opt->heuristic != -1)) {
    dp_set_error("batch_nodes must be >= 1 and heuristic 0 (euclidean) or -1 (blind)");
    return DP_ERR_ARG;
  }
  if (c->P <= 0 || !c->have_world) {
    dp_set_error("upload primitives and world before dp_search");
    return DP_ERR_STATE;
  }
  const int nx = c->desc.nx;
  const dp_dev_sys& S = c->hsys;
  const double t0 = now_ms();
  std::memset(info, 0, sizeof(*info));

  dp_expand_params par;
  std::memset(&par, 0, sizeof(par));
  par.delta = opt->delta;
  par.alpha = opt->alpha;
  par.delta_factor_goal = opt->delta_factor_goal;
  par.cost_delta_factor = opt->cost_delta_factor;
  par.limit_branching_factor = opt->limit_branching_factor;
  par.num_check_goal = 4;  // dbastar.cpp:734-736
  par.check_intermediate_goal = 1;
  par.use_surrogate = opt->use_surrogate;
  par.fp32_rollout = opt->fp32_rollout;
  for (int k = 0; k < nx; ++k) par.goal[k] = goal[k];
  const double r_dup = (1. - (double)opt->alpha) * (double)opt->delta;  // dbastar.cpp:852
  const double goal_tol = opt->delta_factor_goal * (double)opt->delta;

  std::vector<node_rec> nodes;
  std::vector<double> X;  // [node][nx]
  open_list open;
  open.nodes = &nodes;
  node_grid grid;
  grid.dims = S.system == DP_SYS_QUADROTOR ? 3 : 2;
  // cells at least r_dup / w0 wide (the 3x3(x3) block then holds every
  // neighbour); never narrower than 1 um, so that delta = 0 keeps cell
  // coordinates finite
  grid.inv = 1.0 / std::max((r_dup / S.w[0]) * (1.0 + 1e-9), 1e-6);

  auto h_of = [&](const double* x) { return opt->heuristic == 0 ? dp_lower_bound_time(S, x, goal) : 0.0; };
  {
    node_rec s{};
    s.g = 0.0f;
    s.h = (float)h_of(start);
    s.f = s.g + s.h;
    s.parent = -1;
    s.motion = -1;
    s.inter = -1;
    s.in_open = true;
    nodes.push_back(s);
    X.assign(start, start + nx);
    nodes[0].handle = open.push(0);
    grid.add(start, 0);
  }
  int rc = dp_visited_reset(c);
  if (rc) return rc;
  if ((rc = dp_visited_append(c, start, 1))) return rc;

  int32_t status = -1, solution = -1;
  std::vector<int32_t> popped, nbr;
  std::vector<double> xs, gs;
  std::vector<int32_t> pnode, pprim, pgoal;
  std::vector<int8_t> pres;
  std::vector<double> pfin, pg, ph;
  double t_gpu = 0.0;
  for (;;) {
    // stop_search, dbastar.cpp:701-722
    if (info->expands >= opt->max_expands) {
      status = DP_SEARCH_MAX_EXPANDS;
      break;
    }
    if (now_ms() - t0 > opt->search_timelimit_ms) {
      status = DP_SEARCH_MAX_TIME;
      break;
    }
    if (open.empty()) {
      status = DP_SEARCH_EMPTY_QUEUE;
      break;
    }
    popped.clear();
    while ((int32_t)popped.size() < opt->batch_nodes && !open.empty() &&
           info->expands < opt->max_expands) {
      const int32_t b = open.pop();
      nodes[b].in_open = false;
      ++info->closed;
      ++info->expands;
      const double d = dp_distance(S, &X[(size_t)b * nx], goal);
      if (d < goal_tol) {  // dbastar.cpp:794-802
        status = DP_SEARCH_SOLVED;
        solution = b;
        break;
      }
      popped.push_back(b);
    }
    if (status == DP_SEARCH_SOLVED) break;
    if (popped.empty()) continue;

    const int32_t B = (int32_t)popped.size();
    xs.resize((size_t)B * nx);
    gs.resize(B);
    for (int32_t i = 0; i < B; ++i) {
      std::memcpy(&xs[(size_t)i * nx], &X[(size_t)popped[i] * nx], sizeof(double) * nx);
      gs[i] = (double)nodes[popped[i]].g;
    }
    const double tg = now_ms();
    if ((rc = dp_expand_batch(c, xs.data(), gs.data(), B, &par, nullptr))) return rc;
    dp_counts cnt;
    if ((rc = dp_get_counts(c, &cnt))) return rc;
    const int64_t n = cnt.pairs;
    pnode.resize(n);
    pprim.resize(n);
    pgoal.resize(n);
    pres.resize(n);
    pfin.resize((size_t)n * nx);
    pg.resize(n);
    ph.resize(n);
    dp_pairs_out po;
    std::memset(&po, 0, sizeof(po));
    po.capacity = n;
    po.node = pnode.data();
    po.prim = pprim.data();
    po.goal_index = pgoal.data();
    po.resolution = pres.data();
    po.final_state = pfin.data();
    po.g = pg.data();
    po.h = ph.data();
    if (n > 0 && (rc = dp_fetch_pairs(c, &po))) return rc;
    if ((rc = dp_visited_append_novel(c))) return rc;
    t_gpu += now_ms() - tg;
    ++info->gpu_batches;
    info->pairs += n;
    info->checked += cnt.checked;

    for (int64_t q = 0; q < n; ++q) {
      const int8_t r = pres[q];
      if (r != DP_R_NOVEL && r != DP_R_DUP_VISITED && r != DP_R_DUP_BATCH) continue;
      const int32_t parent = popped[pnode[q]];
      const double* x = &pfin[(size_t)q * nx];
      const double gq = pg[q];
      if (r == DP_R_NOVEL) {  // dbastar.cpp:856-876
        const double hq = opt->heuristic == 0 ? ph[q] : 0.0;
        node_rec m{};
        m.g = (float)gq;
        m.h = (float)hq;
        m.f = (float)(gq + hq);
        m.parent = parent;
        m.motion = pprim[q];
        m.inter = pgoal[q];
        m.in_open = true;
        const int32_t id = (int32_t)nodes.size();
        nodes.push_back(m);
        X.insert(X.end(), x, x + nx);
        nodes[id].handle = open.push(id);
        grid.add(x, id);
        continue;
      }
      // rewiring of every neighbour within (1 - alpha) delta, dbastar.cpp:883-910,
      // in ascending node order
      nbr.clear();
      grid.around(x, [&](int32_t id) {
        if (dp_distance(S, x, &X[(size_t)id * nx]) <= r_dup) nbr.push_back(id);
      });
      std::sort(nbr.begin(), nbr.end());
      for (int32_t id : nbr) {
        const double* y = &X[(size_t)id * nx];
        const double tentative = gq + dp_lower_bound_time(S, x, y);
        node_rec& m = nodes[id];
        if (tentative < m.g) {
          m.g = (float)tentative;
          m.f = (float)(tentative + m.h);
          m.parent = parent;
          m.motion = pprim[q];
          m.inter = -1;
          ++info->rewires;
          if (m.in_open)
            open.increase(m.handle);
          else
            m.handle = open.push(id);
        }
      }
    }
  }

  if (status != DP_SEARCH_SOLVED) {  // closest node to the goal (dbastar.cpp:944-952)
    double best = INFINITY;
    for (size_t i = 0; i < nodes.size(); ++i) {
      const double d = dp_distance(S, &X[i * nx], goal);
      if (d < best) {
        best = d;
        solution = (int32_t)i;
      }
    }
  }
  info->status = status;
  info->solved = status == DP_SEARCH_SOLVED;
  info->cost = nodes[solution].g;
  info->distance_to_goal = dp_distance(S, &X[(size_t)solution * nx], goal);
  info->all_nodes = (int64_t)nodes.size();

  // from_solution_to_yaml_and_traj, dbastar.cpp:129-310
  std::vector<int32_t> path;
  for (int32_t v = solution; v != -1; v = nodes[v].parent) path.push_back(v);
  std::reverse(path.begin(), path.end());
  c->sol_states.clear();
  c->sol_actions.clear();
  if (path.size() == 1) c->sol_states.insert(c->sol_states.end(), &X[(size_t)path[0] * nx], &X[(size_t)path[0] * nx] + nx);
  std::vector<double> traj;
  const int nu = c->h_nu;
  for (size_t i = 0; i + 1 < path.size(); ++i) {
    const node_rec& child = nodes[path[i + 1]];
    transform(c, &X[(size_t)path[i] * nx], child.motion, traj);
    const int64_t L1 = (int64_t)traj.size() / nx;
    const int64_t take = child.inter != -1 ? child.inter + 1 : L1;
    for (int64_t k = 0; k < take; ++k) {
      if (k < take - 1)
        c->sol_states.insert(c->sol_states.end(), &traj[(size_t)k * nx], &traj[(size_t)k * nx] + nx);
      else if (i + 2 == path.size())
        c->sol_states.insert(c->sol_states.end(), &X[(size_t)path[i + 1] * nx], &X[(size_t)path[i + 1] * nx] + nx);
    }
    if (nu) {
      const int64_t na = child.inter != -1 ? child.inter : L1 - 1;
      const int32_t a0 = c->h_off[child.motion] - child.motion;
      for (int64_t k = 0; k < na; ++k)
        c->sol_actions.insert(c->sol_actions.end(), &c->h_ac[(size_t)(a0 + k) * nu],
                              &c->h_ac[(size_t)(a0 + k) * nu] + nu);
    }
  }
  c->sol_nu = nu;
  info->traj_states = (int64_t)(c->sol_states.size() / nx);
  info->traj_actions = nu ? (int64_t)(c->sol_actions.size() / nu) : 0;
  info->time_gpu_ms = t_gpu;
  info->time_search_ms = now_ms() - t0;
  return DP_OK;
}

int dp_search_trajectory(dp_ctx* c, double* states, double* actions) {
  if (!c) {
    dp_set_error("null context");
    return DP_ERR_ARG;
  }
  if (states && !c->sol_states.empty()) std::memcpy(states, c->sol_states.data(), sizeof(double) * c->sol_states.size());
  if (actions && !c->sol_actions.empty())
    std::memcpy(actions, c->sol_actions.data(), sizeof(double) * c->sol_actions.size());
  return DP_OK;
}

int dp_search_trajectory_size(const dp_ctx* c, int64_t* n_states, int64_t* n_actions) {
  if (!c || !n_states || !n_actions) {
    dp_set_error("null argument");
    return DP_ERR_ARG;
  }
  *n_states = (int64_t)(c->sol_states.size() / c->desc.nx);
  *n_actions = c->sol_nu ? (int64_t)(c->sol_actions.size() / c->sol_nu) : 0;
  return DP_OK;
}

void dp_idbastar_options_default(dp_idbastar_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->delta_0 = 0.3;  // idbastar.hpp:48-59
  o->num_primitives_0 = 1000;
  o->delta_rate = 0.9;
  o->num_primitives_rate = 1.5;
  o->timelimit_s = 10.0;
  o->max_it = 10;
  o->max_num_sol = 5;
  o->add_primitives = 1;
  o->cuts = 5;
  o->seed = 1;
}

int dp_idbastar(dp_ctx* c, const dp_prims* lib, const double* start, const double* goal,
                const dp_search_options* sopt, const dp_idbastar_options* o, dp_idbastar_info* info) {
  if (!c || !lib || !lib->offsets || !lib->states || !start || !goal || !sopt || !o || !info || lib->P <= 0 ||
      o->cuts < 1 || o->num_primitives_0 < 1) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  const int nx = c->desc.nx;
  const int nu = lib->actions ? lib->nu : 0;
  if (lib->nx != nx) {
    dp_set_error("primitive nx does not match the system");
    return DP_ERR_ARG;
  }
  std::memset(info, 0, sizeof(*info));
  info->cost = 1e8;
  const double t0 = now_ms();
  // motions inserted from solutions, newest first; each is states [L1][nx] + actions [L1-1][nu]
  std::vector<std::vector<double>> add_st, add_ac;
  std::vector<double> best_states, best_actions;
  uint64_t rng = o->seed ? o->seed : 1;
  auto unif = [&]() {  // U[-1, 1] from a splitmix64 stream
    uint64_t z = (rng += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
  };
  const dp_dev_sys& S = c->hsys;
  // Options_dbastar::delta is a float and max_motions a size_t: each
  // update of idbastar.cpp:128-146 rounds (truncates) to them
  float delta = (float)o->delta_0;
  int64_t max_motions = o->num_primitives_0;
  bool solved_db = false;
  int rc;
  for (int32_t it = 0;; ++it) {
    if (it > 0) {  // idbastar.cpp:128-146 (new_schedule)
      delta = (float)((double)delta * (solved_db ? o->delta_rate : 0.9999));
      max_motions = (int64_t)((double)max_motions * o->num_primitives_rate);
    }
    // primitive set of this iteration: inserted motions, then the library
    std::vector<int32_t> off{0};
    std::vector<double> st, ac;
    for (size_t m = 0; m < add_st.size(); ++m) {
      st.insert(st.end(), add_st[m].begin(), add_st[m].end());
      ac.insert(ac.end(), add_ac[m].begin(), add_ac[m].end());
      off.push_back(off.back() + (int32_t)(add_st[m].size() / nx));
    }
    for (int32_t j = 0; j < lib->P; ++j) {
      const int32_t a = lib->offsets[j], b = lib->offsets[j + 1];
      st.insert(st.end(), lib->states + (int64_t)a * nx, lib->states + (int64_t)b * nx);
      if (nu) ac.insert(ac.end(), lib->actions + (int64_t)(a - j) * nu, lib->actions + (int64_t)(b - j - 1) * nu);
      off.push_back(off.back() + (b - a));
    }
    dp_prims pr;
    std::memset(&pr, 0, sizeof(pr));
    pr.P = (int32_t)off.size() - 1;
    pr.nx = nx;
    pr.nu = nu;
    pr.offsets = off.data();
    pr.states = st.data();
    pr.actions = nu ? ac.data() : nullptr;
    const int32_t use = (int32_t)std::min<int64_t>(max_motions, (int64_t)pr.P);
    if ((rc = dp_upload_primitives(c, &pr, use))) return rc;

    dp_search_options so = *sopt;
    so.delta = delta;
    dp_search_info si;
    if ((rc = dp_search(c, start, goal, &so, &si))) return rc;
    solved_db = si.solved != 0;
    info->iterations = it + 1;
    info->last_delta = delta;
    info->last_primitives = use;
    info->expands += si.expands;
    if (solved_db) {
      ++info->solutions;
      info->solved = 1;
      if (si.cost < info->cost) {
        info->cost = si.cost;
        best_states = c->sol_states;
        best_actions = c->sol_actions;
      }
      const int64_t ns = (int64_t)(c->sol_states.size() / nx);
      if (o->add_primitives && nu && ns >= 2) {
        // cut into pieces of about equal numbers of actions, canonical, noise
        const int64_t na = ns - 1;
        const int32_t cuts = (int32_t)std::min<int64_t>(o->cuts, na);
        std::vector<std::vector<double>> pst, pac;
        for (int32_t k = 0; k < cuts; ++k) {
          const int64_t a0 = na * k / cuts, a1 = na * (k + 1) / cuts;
          std::vector<double> ps(c->sol_states.begin() + a0 * nx, c->sol_states.begin() + (a1 + 1) * nx);
          std::vector<double> pa(c->sol_actions.begin() + a0 * nu, c->sol_actions.begin() + a1 * nu);
          double off0[DP_MAX_NX];
          for (int d = 0; d < nx; ++d) off0[d] = ps[d];
          for (int64_t r = 0; r <= a1 - a0; ++r)
            for (int d = 0; d < nx; ++d)
              if (dp_is_offset_dim(S, d)) ps[(size_t)r * nx + d] -= off0[d];
          for (int64_t r : {(int64_t)0, a1 - a0}) {
            for (int d = 0; d < nx; ++d) ps[(size_t)r * nx + d] += 1e-7 * unif();
            if (S.system == DP_SYS_QUADROTOR) {  // keep the quaternion unit
              double* q = &ps[(size_t)r * nx + 3];
              const double n = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
              for (int e = 0; e < 4; ++e) q[e] /= n;
            }
          }
          pst.push_back(std::move(ps));
          pac.push_back(std::move(pa));
        }
        add_st.insert(add_st.begin(), pst.begin(), pst.end());
        add_ac.insert(add_ac.begin(), pac.begin(), pac.end());
        info->primitives_added += cuts;
      }
    }
    // exit criteria, idbastar.cpp:300-313 (a later one overrides an earlier)
    bool finished = false;
    if (it + 1 >= o->max_it) {
      finished = true;
      info->exit_criteria = DP_IDB_MAX_IT;
    }
    if (info->solutions >= o->max_num_sol) {
      finished = true;
      info->exit_criteria = DP_IDB_MAX_SOLUTION;
    }
    if ((now_ms() - t0) / 1000.0 > o->timelimit_s) {
      finished = true;
      info->exit_criteria = DP_IDB_TIME_LIMIT;
    }
    if (finished) break;
  }
  if (info->solved) {
    c->sol_states = best_states;
    c->sol_actions = best_actions;
  }
  info->time_ms = now_ms() - t0;
  return DP_OK;
}

}  // extern "C"
