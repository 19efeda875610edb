// Host-side HIN utilities: seeded synthetic generators, reference text-format IO and the
// meta-path parser.  See include/smcm_hin.h for the contract.
//
// Generators: RNG is xoshiro256** seeded through splitmix64; finite Zipf is sampled by
// inverse CDF (binary search over the cumulative weights k^-s); Poisson by Knuth's
// multiplication method; Geometric (failures before first success, support >= 0) by
// inversion; "distinct" draws use rejection.  Each relation draws from its own stream
// (seed, relation id) so that adding a relation never perturbs the others.
#include "smcm_hin.h"

#include <algorithm>
#include <functional>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

struct Rng {
  uint64_t s[4];
  static uint64_t splitmix(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  Rng(uint64_t seed, uint64_t stream) {
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ull);
    for (auto& w : s) w = splitmix(x);
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
  // uniform integer in [0, n)
  int32_t below(int32_t n) { return (int32_t)(uniform() * (double)n); }
};

// Finite Zipf over ranks 1..n with exponent s, mapped to ids through a seeded permutation so
// that hubs are scattered over the id range.
struct Zipf {
  std::vector<double> cdf;
  std::vector<int32_t> perm;
  Zipf(int32_t n, double s, Rng& perm_rng) : cdf(n), perm(n) {
    double acc = 0.0;
    for (int32_t k = 0; k < n; ++k) {
      acc += std::pow((double)(k + 1), -s);
      cdf[k] = acc;
    }
    for (int32_t k = 0; k < n; ++k) cdf[k] /= acc;
    cdf[n - 1] = 1.0;
    for (int32_t k = 0; k < n; ++k) perm[k] = k;
    for (int32_t k = n - 1; k > 0; --k) {
      int32_t j = (int32_t)(perm_rng.uniform() * (double)(k + 1));
      std::swap(perm[k], perm[j]);
    }
  }
  int32_t draw(Rng& r) const {
    double u = r.uniform();
    int32_t k = (int32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
    if (k >= (int32_t)cdf.size()) k = (int32_t)cdf.size() - 1;
    return perm[k];
  }
};

int32_t poisson(Rng& r, double lam) {
  double L = std::exp(-lam), p = 1.0;
  int32_t k = 0;
  do {
    ++k;
    p *= r.uniform();
  } while (p > L);
  return k - 1;
}

int32_t geometric(Rng& r, double p) {
  double u = r.uniform();
  if (u <= 0.0) u = 1e-300;
  return (int32_t)std::floor(std::log(u) / std::log1p(-p));
}

// draw `k` distinct values with `draw`, returned ascending
template <class F>
void draw_distinct(int32_t k, int32_t universe, F draw, std::vector<int32_t>& out) {
  out.clear();
  if (k > universe) k = universe;
  if (k <= 64) {
    while ((int32_t)out.size() < k) {
      int32_t x = draw();
      if (std::find(out.begin(), out.end(), x) == out.end()) out.push_back(x);
    }
  } else {
    std::unordered_set<int32_t> seen;
    while ((int32_t)out.size() < k) {
      int32_t x = draw();
      if (seen.insert(x).second) out.push_back(x);
    }
  }
  std::sort(out.begin(), out.end());
}

hin_graph* alloc_graph(int32_t n, int64_t m) {
  hin_graph* g = (hin_graph*)std::calloc(1, sizeof(hin_graph));
  g->n = n;
  g->m = m;
  g->vtype = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max(n, 1));
  g->adj_ptr = (int64_t*)std::malloc(sizeof(int64_t) * (size_t)(n + 1));
  g->adj_nbr = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
  g->adj_et = (int32_t*)std::malloc(sizeof(int32_t) * (size_t)std::max<int64_t>(m, 1));
  return g;
}

void finish_meta(hin_graph* g) {
  int32_t maxt = -1;
  for (int32_t v = 0; v < g->n; ++v) maxt = std::max(maxt, g->vtype[v]);
  g->ntypes = maxt + 1;
  for (int t = 0; t < HIN_MAX_TYPES; ++t) {
    g->dims[t] = 0;
    g->type_offset[t] = -1;
  }
  std::vector<int32_t> first(HIN_MAX_TYPES, -1), last(HIN_MAX_TYPES, -1);
  for (int32_t v = 0; v < g->n; ++v) {
    int32_t t = g->vtype[v];
    if (t < 0 || t >= HIN_MAX_TYPES) continue;
    g->dims[t]++;
    if (first[t] < 0) first[t] = v;
    last[t] = v;
  }
  for (int t = 0; t < HIN_MAX_TYPES; ++t)
    if (g->dims[t] > 0 && last[t] - first[t] + 1 == g->dims[t]) g->type_offset[t] = first[t];
}

// Assemble a graph from per-vertex neighbour lists given as (src, dst, et) triples grouped
// by relation.  Each vertex's neighbour list is sorted ascending by neighbour id (stable on
// insertion order for equal ids).
struct Triple {
  int32_t src, dst, et;
};

hin_graph* assemble(int32_t n, const std::vector<int32_t>& vtype, std::vector<Triple>& edges) {
  std::stable_sort(edges.begin(), edges.end(), [](const Triple& a, const Triple& b) {
    return a.src != b.src ? a.src < b.src : a.dst < b.dst;
  });
  hin_graph* g = alloc_graph(n, (int64_t)edges.size());
  std::memcpy(g->vtype, vtype.data(), sizeof(int32_t) * n);
  std::fill(g->adj_ptr, g->adj_ptr + n + 1, 0);
  for (auto& e : edges) g->adj_ptr[e.src + 1]++;
  for (int32_t v = 0; v < n; ++v) g->adj_ptr[v + 1] += g->adj_ptr[v];
  for (size_t i = 0; i < edges.size(); ++i) {
    g->adj_nbr[i] = edges[i].dst;
    g->adj_et[i] = edges[i].et;
  }
  finish_meta(g);
  return g;
}

}  // namespace

extern "C" {

const char* hin_last_error(void) { return g_err.c_str(); }

int hin_gen_bib(int32_t nA, int32_t nP, int32_t nV, int32_t nT, double lam_auth, double s_auth,
                double s_venue, int32_t terms_per_paper, double s_term, uint64_t seed,
                hin_graph** out) {
  if (!out || nA <= 0 || nP <= 0 || nV <= 0 || nT < 0) return fail(-EINVAL, "hin_gen_bib: bad sizes");
  const int32_t offP = 0, offA = nP, offV = nP + nA, offT = nP + nA + nV;
  const int32_t n = nP + nA + nV + nT;
  std::vector<int32_t> vtype(n);
  for (int32_t i = 0; i < nP; ++i) vtype[offP + i] = 0;
  for (int32_t i = 0; i < nA; ++i) vtype[offA + i] = 1;
  for (int32_t i = 0; i < nV; ++i) vtype[offV + i] = 2;
  for (int32_t i = 0; i < nT; ++i) vtype[offT + i] = 3;

  Rng perm_rng(seed, 100);
  Zipf zA(nA, s_auth, perm_rng), zV(nV, s_venue, perm_rng);
  Rng rA(seed, 1), rV(seed, 2), rT(seed, 3);
  std::vector<Triple> edges;
  edges.reserve((size_t)nP * (size_t)(2 * (2 + (int)lam_auth) + 2 + 2 * terms_per_paper));
  std::vector<int32_t> pick;
  for (int32_t p = 0; p < nP; ++p) {
    int32_t k = 1 + poisson(rA, lam_auth);
    draw_distinct(k, nA, [&] { return zA.draw(rA); }, pick);
    for (int32_t a : pick) {
      edges.push_back({offP + p, offA + a, 0});
      edges.push_back({offA + a, offP + p, 3});
    }
    int32_t v = zV.draw(rV);
    edges.push_back({offP + p, offV + v, 1});
    edges.push_back({offV + v, offP + p, 4});
  }
  if (nT > 0 && terms_per_paper > 0) {
    Rng perm_t(seed, 101);
    Zipf zT(nT, s_term, perm_t);
    for (int32_t p = 0; p < nP; ++p) {
      draw_distinct(terms_per_paper, nT, [&] { return zT.draw(rT); }, pick);
      for (int32_t t : pick) {
        edges.push_back({offP + p, offT + t, 2});
        edges.push_back({offT + t, offP + p, 5});
      }
    }
  }
  *out = assemble(n, vtype, edges);
  return 0;
}

int hin_gen_loc(int32_t nU, int32_t nL, int32_t nC, int32_t nY, int32_t nS, double geo_p,
                double s_checkin, double s_city, uint64_t seed, hin_graph** out) {
  if (!out || nU <= 0 || nL <= 0 || nC <= 0 || nY <= 0 || nS <= 0)
    return fail(-EINVAL, "hin_gen_loc: bad sizes");
  // type order by type id: venue 0, city 1, category 2, user 3, slot 4
  const int32_t offL = 0, offY = nL, offC = nL + nY, offU = nL + nY + nC, offS = nL + nY + nC + nU;
  const int32_t n = nL + nY + nC + nU + nS;
  std::vector<int32_t> vtype(n);
  for (int32_t i = 0; i < nL; ++i) vtype[offL + i] = 0;
  for (int32_t i = 0; i < nY; ++i) vtype[offY + i] = 1;
  for (int32_t i = 0; i < nC; ++i) vtype[offC + i] = 2;
  for (int32_t i = 0; i < nU; ++i) vtype[offU + i] = 3;
  for (int32_t i = 0; i < nS; ++i) vtype[offS + i] = 4;

  Rng perm_rng(seed, 100);
  Zipf zL(nL, s_checkin, perm_rng), zY(nY, s_city, perm_rng);
  Rng rU(seed, 1), rY(seed, 2), rC(seed, 3), rS(seed, 4);
  std::vector<Triple> edges;
  std::vector<int32_t> pick;
  for (int32_t u = 0; u < nU; ++u) {
    int32_t k = 1 + geometric(rU, geo_p);
    draw_distinct(k, nL, [&] { return zL.draw(rU); }, pick);
    for (int32_t l : pick) {
      edges.push_back({offU + u, offL + l, 5});
      edges.push_back({offL + l, offU + u, 4});
    }
  }
  for (int32_t l = 0; l < nL; ++l) {
    int32_t y = zY.draw(rY);
    edges.push_back({offL + l, offY + y, 0});
    edges.push_back({offY + y, offL + l, 1});
    int32_t kc = 1 + rC.below(3);
    draw_distinct(kc, nC, [&] { return rC.below(nC); }, pick);
    for (int32_t c : pick) {
      edges.push_back({offL + l, offC + c, 2});
      edges.push_back({offC + c, offL + l, 3});
    }
    int32_t ks = 1 + rS.below(4);
    draw_distinct(ks, nS, [&] { return rS.below(nS); }, pick);
    for (int32_t s : pick) {
      edges.push_back({offL + l, offS + s, 6});
      edges.push_back({offS + s, offL + l, 7});
    }
  }
  *out = assemble(n, vtype, edges);
  return 0;
}

int hin_from_arrays(int32_t n, const int32_t* vtype, const int64_t* adj_ptr,
                    const int32_t* adj_nbr, const int32_t* adj_et, hin_graph** out) {
  if (!out || n < 0 || !vtype || !adj_ptr) return fail(-EINVAL, "hin_from_arrays: null input");
  int64_t m = adj_ptr[n];
  if (adj_ptr[0] != 0 || m < 0) return fail(-EINVAL, "hin_from_arrays: bad adj_ptr");
  for (int32_t v = 0; v < n; ++v)
    if (adj_ptr[v + 1] < adj_ptr[v]) return fail(-EINVAL, "hin_from_arrays: adj_ptr not monotone");
  for (int32_t v = 0; v < n; ++v)
    if (vtype[v] < 0 || vtype[v] >= HIN_MAX_TYPES) return fail(-EINVAL, "hin_from_arrays: vertex type out of range");
  for (int64_t e = 0; e < m; ++e)
    if (adj_nbr[e] < 0 || adj_nbr[e] >= n) return fail(-EINVAL, "hin_from_arrays: neighbour id out of range");
  hin_graph* g = alloc_graph(n, m);
  std::memcpy(g->vtype, vtype, sizeof(int32_t) * n);
  std::memcpy(g->adj_ptr, adj_ptr, sizeof(int64_t) * (n + 1));
  if (m) {
    std::memcpy(g->adj_nbr, adj_nbr, sizeof(int32_t) * m);
    std::memcpy(g->adj_et, adj_et, sizeof(int32_t) * m);
  }
  finish_meta(g);
  *out = g;
  return 0;
}

int hin_write_text(const hin_graph* g, const char* graph_file, const char* vertex_file,
                   const char* edge_file) {
  if (!g) return fail(-EINVAL, "hin_write_text: null graph");
  FILE* fv = std::fopen(vertex_file, "w");
  FILE* fe = std::fopen(edge_file, "w");
  FILE* fg = std::fopen(graph_file, "w");
  if (!fv || !fe || !fg) {
    if (fv) std::fclose(fv);
    if (fe) std::fclose(fe);
    if (fg) std::fclose(fg);
    return fail(-EIO, "hin_write_text: cannot open output files");
  }
  for (int32_t v = 0; v < g->n; ++v) std::fprintf(fv, "%d %d\n", v, g->vtype[v]);
  for (int64_t e = 0; e < g->m; ++e) std::fprintf(fe, "%lld %d\n", (long long)e, g->adj_et[e]);
  for (int32_t v = 0; v < g->n; ++v) {
    std::fprintf(fg, "%d", v);
    for (int64_t e = g->adj_ptr[v]; e < g->adj_ptr[v + 1]; ++e)
      std::fprintf(fg, " %d %lld", g->adj_nbr[e], (long long)e);
    std::fputc('\n', fg);
  }
  std::fclose(fv);
  std::fclose(fe);
  std::fclose(fg);
  return 0;
}

// Loader with the reference's semantics (efficient/resmcm.cpp:59-176): n = number of lines of
// vertex.txt; vertexTypes[v] from "v vt" lines; edge types from "e et" lines; graph.txt lines
// "v nb e nb e ..." resolve each entry's type through the edge-type table, in file order.
// Vertices without a graph.txt line get an empty neighbour list (the reference leaves such
// rows uninitialised, resmcm.cpp:87,105; here they are defined as empty).
int hin_read_text(const char* graph_file, const char* vertex_file, const char* edge_file,
                  hin_graph** out) {
  if (!out) return fail(-EINVAL, "hin_read_text: null out");
  std::ifstream fv(vertex_file);
  if (!fv) return fail(-EIO, std::string("hin_read_text: can not open ") + vertex_file);
  std::vector<std::pair<int64_t, int32_t>> vlines;
  std::string line;
  while (std::getline(fv, line)) {
    std::istringstream b(line);
    long long v = -1;
    int vt = 0;
    b >> v >> vt;
    vlines.push_back({v, vt});
  }
  const int64_t n64 = (int64_t)vlines.size();
  if (n64 > INT32_MAX) return fail(-EINVAL, "hin_read_text: too many vertices");
  const int32_t n = (int32_t)n64;
  std::vector<int32_t> vtype(n, 0);
  for (auto& p : vlines) {
    if (p.first < 0 || p.first >= n) return fail(-EINVAL, "hin_read_text: vertex id out of range");
    if (p.second < 0 || p.second >= HIN_MAX_TYPES) return fail(-EINVAL, "hin_read_text: vertex type out of range");
    vtype[p.first] = p.second;
  }
  std::vector<int32_t> etype;
  {
    std::ifstream fe(edge_file);
    if (!fe) return fail(-EIO, std::string("hin_read_text: can not open ") + edge_file);
    while (std::getline(fe, line)) {
      std::istringstream b(line);
      long long e = -1;
      int et = 0;
      b >> e >> et;
      if (e < 0) continue;
      if ((size_t)e >= etype.size()) etype.resize((size_t)e + 1, -1);
      etype[e] = et;
    }
  }
  std::ifstream fg(graph_file);
  if (!fg) return fail(-EIO, std::string("hin_read_text: can not open ") + graph_file);
  std::vector<std::vector<std::pair<int32_t, int32_t>>> rows(n);
  while (std::getline(fg, line)) {
    std::istringstream b(line);
    long long v = -1;
    if (!(b >> v)) continue;
    if (v < 0 || v >= n) return fail(-EINVAL, "hin_read_text: graph vertex id out of range");
    auto& r = rows[v];
    r.clear();  // a later line for the same vertex replaces the earlier one (resmcm.cpp:105)
    long long nb, e;
    while (b >> nb) {
      if (!(b >> e)) break;
      if (nb < 0 || nb >= n) return fail(-EINVAL, "hin_read_text: neighbour id out of range");
      if (e < 0 || (size_t)e >= etype.size()) return fail(-EINVAL, "hin_read_text: edge id without a type");
      r.push_back({(int32_t)nb, etype[e]});
    }
  }
  int64_t m = 0;
  for (auto& r : rows) m += (int64_t)r.size();
  hin_graph* g = alloc_graph(n, m);
  std::memcpy(g->vtype, vtype.data(), sizeof(int32_t) * n);
  g->adj_ptr[0] = 0;
  int64_t k = 0;
  for (int32_t v = 0; v < n; ++v) {
    for (auto& p : rows[v]) {
      g->adj_nbr[k] = p.first;
      g->adj_et[k] = p.second;
      ++k;
    }
    g->adj_ptr[v + 1] = k;
  }
  finish_meta(g);
  *out = g;
  return 0;
}

void hin_free(hin_graph* g) {
  if (!g) return;
  std::free(g->vtype);
  std::free(g->adj_ptr);
  std::free(g->adj_nbr);
  std::free(g->adj_et);
  std::free(g);
}

int hin_parse_metapath(const char* s, int32_t* vtypes, int32_t* etypes, int32_t cap, int32_t* L) {
  if (!s || !L) return fail(-EINVAL, "hin_parse_metapath: null input");
  std::vector<int32_t> vt, et;
  size_t len = std::strlen(s);
  if (len == 0) return fail(-EINVAL, "the meta-path is incorrect");
  int64_t num = 0;
  int i = 0;
  for (size_t c = 0; c < len; ++c) {
    char ch = s[c];
    if (ch >= '0' && ch <= '9') {
      num = num * 10 + (ch - '0');
      if (num > INT32_MAX) return fail(-EINVAL, "hin_parse_metapath: number too large");
    } else {
      if (i % 2 == 0) vt.push_back((int32_t)num);
      else et.push_back((int32_t)num);
      num = 0;
      ++i;
    }
  }
  char last = s[len - 1];
  if (last >= '0' && last <= '9') {
    if (i % 2 == 0) vt.push_back((int32_t)num);
    else et.push_back((int32_t)num);
    ++i;
  }
  if (vt.size() != et.size() + 1) return fail(-EINVAL, "the meta-path is incorrect");
  int32_t nmat = i / 2;
  if ((int32_t)vt.size() > cap) return fail(-ERANGE, "hin_parse_metapath: capacity too small");
  for (size_t k = 0; k < vt.size(); ++k) if (vtypes) vtypes[k] = vt[k];
  for (size_t k = 0; k < et.size(); ++k) if (etypes) etypes[k] = et[k];
  *L = nmat;
  return 0;
}

int hin_enumerate_metapaths(const int32_t* rel_src, const int32_t* rel_et, const int32_t* rel_dst,
                            int32_t nrel, const int32_t* starts, int32_t nstarts, int32_t L,
                            int32_t* out_vts, int32_t* out_ets, int64_t cap, int64_t* count) {
  if (!count || nrel < 0 || nstarts < 0 || L < 0 || (nrel > 0 && (!rel_src || !rel_et || !rel_dst)) ||
      (nstarts > 0 && !starts) || (cap > 0 && (!out_vts || (L > 0 && !out_ets)))) {
    g_err = "bad arguments";
    return -22;
  }
  int64_t n = 0;
  std::vector<int32_t> vts(L + 1), ets(L > 0 ? L : 1);
  // depth-first, neighbours in relation order (recursion depth is L)
  std::function<void(int32_t)> dfs = [&](int32_t depth) {
    if (depth == L) {
      if (n < cap) {
        std::copy(vts.begin(), vts.end(), out_vts + n * (L + 1));
        if (L > 0) std::copy(ets.begin(), ets.begin() + L, out_ets + n * L);
      }
      ++n;
      return;
    }
    for (int32_t r = 0; r < nrel; ++r) {
      if (rel_src[r] != vts[depth]) continue;
      ets[depth] = rel_et[r];
      vts[depth + 1] = rel_dst[r];
      dfs(depth + 1);
    }
  };
  for (int32_t i = 0; i < nstarts; ++i) {
    vts[0] = starts[i];
    dfs(0);
  }
  *count = n;
  return 0;
}

int hin_metapath_to_string(const int32_t* vtypes, const int32_t* etypes, int32_t L, char* buf,
                           int32_t cap) {
  std::string s;
  for (int32_t i = 0; i < L; ++i) s += std::to_string(vtypes[i]) + "-" + std::to_string(etypes[i]) + "-";
  s += std::to_string(vtypes[L]);
  if (buf && cap > 0) {
    std::strncpy(buf, s.c_str(), (size_t)cap - 1);
    buf[cap - 1] = 0;
  }
  return (int)s.size();
}

}  // extern "C"
