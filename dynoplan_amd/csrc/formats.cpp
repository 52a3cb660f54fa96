// formats.cpp -- motion-primitive files (the on-disk boundary of the path).
//
// load_motion_primitives_new (src/ompl/robots.cpp:3061-3174) restated on
// plain arrays:
//   1. format by extension (.yaml .json .msgpack .bin), robots.cpp:3070-3080
//   2. dynobench::Trajectories::load_file_{yaml,json,msgpack,boost}
//   3. truncate to max_motions                             :3104-3105
//   4. endpoint noise: first and last state += noise * U[-1,1]  :3113-3127
//      (the reference draws from an unseeded std::rand; here the draw is a
//       seeded splitmix64 stream, and noise = 0 disables it)
//   5. quadrotor: normalise the quaternion of every state   :3130-3137
//   6. cost > 1e5 on the first motion -> every cost = ref_dt * #actions
//                                                           :3153-3159
// Files hold a list of trajectories; each has "states" (list of state
// vectors) and "actions" (list of control vectors), optionally "cost".
// msgpack / JSON wrap the list as {"data": [...]}; YAML is either the bare
// list or {data: [...]}.  The Boost binary archive (binary_oarchive, library
// version 17) stores per trajectory: states, actions (count u64, item version
// u32, then size-prefixed f64 vectors), cost f64, feasible u8, start and goal
// vectors, info string; the first trajectory also carries class records.
// The writers produce the same layouts, so files round-trip bit-exactly.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/dynoplan_amd.h"

void dp_set_error(const std::string& msg);

namespace {

// ---- a small document tree -------------------------------------------------------
struct Node {
  enum Kind { NIL, BOOL, NUM, STR, ARR, MAP } kind = NIL;
  double num = 0.0;
  std::string str;
  std::vector<Node> arr;
  std::vector<std::pair<std::string, Node>> map;
  const Node* get(const std::string& k) const {
    for (auto& kv : map)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

struct Traj {
  std::vector<std::vector<double>> states, actions;
  double cost = 1e8;  // dynobench's default for an unknown cost
};

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---- msgpack ---------------------------------------------------------------------
struct MsgpackReader {
  const uint8_t* p;
  const uint8_t* e;
  void need(size_t n) {
    if ((size_t)(e - p) < n) throw ParseError("msgpack: truncated");
  }
  uint64_t be(int n) {
    need(n);
    uint64_t v = 0;
    for (int i = 0; i < n; ++i) v = (v << 8) | p[i];
    p += n;
    return v;
  }
  Node value(int depth = 0) {
    if (depth > 64) throw ParseError("msgpack: nesting too deep");
    need(1);
    const uint8_t t = *p++;
    Node n;
    auto array = [&](uint64_t len) {
      n.kind = Node::ARR;
      n.arr.reserve(len);
      for (uint64_t i = 0; i < len; ++i) n.arr.push_back(value(depth + 1));
    };
    auto map = [&](uint64_t len) {
      n.kind = Node::MAP;
      for (uint64_t i = 0; i < len; ++i) {
        Node k = value(depth + 1);
        Node v = value(depth + 1);
        n.map.emplace_back(k.kind == Node::STR ? k.str : std::to_string(k.num), std::move(v));
      }
    };
    auto str = [&](uint64_t len) {
      need(len);
      n.kind = Node::STR;
      n.str.assign((const char*)p, len);
      p += len;
    };
    if (t <= 0x7f) { n.kind = Node::NUM; n.num = t; }
    else if (t >= 0xe0) { n.kind = Node::NUM; n.num = (int8_t)t; }
    else if ((t & 0xf0) == 0x80) map(t & 0x0f);
    else if ((t & 0xf0) == 0x90) array(t & 0x0f);
    else if ((t & 0xe0) == 0xa0) str(t & 0x1f);
    else switch (t) {
      case 0xc0: n.kind = Node::NIL; break;
      case 0xc2: n.kind = Node::BOOL; n.num = 0; break;
      case 0xc3: n.kind = Node::BOOL; n.num = 1; break;
      case 0xc4: str(be(1)); break;
      case 0xc5: str(be(2)); break;
      case 0xc6: str(be(4)); break;
      case 0xca: { uint32_t u = (uint32_t)be(4); float f; std::memcpy(&f, &u, 4); n.kind = Node::NUM; n.num = f; break; }
      case 0xcb: { uint64_t u = be(8); double d; std::memcpy(&d, &u, 8); n.kind = Node::NUM; n.num = d; break; }
      case 0xcc: n.kind = Node::NUM; n.num = (double)be(1); break;
      case 0xcd: n.kind = Node::NUM; n.num = (double)be(2); break;
      case 0xce: n.kind = Node::NUM; n.num = (double)be(4); break;
      case 0xcf: n.kind = Node::NUM; n.num = (double)be(8); break;
      case 0xd0: n.kind = Node::NUM; n.num = (int8_t)be(1); break;
      case 0xd1: n.kind = Node::NUM; n.num = (int16_t)be(2); break;
      case 0xd2: n.kind = Node::NUM; n.num = (int32_t)be(4); break;
      case 0xd3: n.kind = Node::NUM; n.num = (double)(int64_t)be(8); break;
      case 0xd9: str(be(1)); break;
      case 0xda: str(be(2)); break;
      case 0xdb: str(be(4)); break;
      case 0xdc: array(be(2)); break;
      case 0xdd: array(be(4)); break;
      case 0xde: map(be(2)); break;
      case 0xdf: map(be(4)); break;
      default: throw ParseError("msgpack: unsupported type byte");
    }
    return n;
  }
};

// ---- JSON ------------------------------------------------------------------------
struct JsonReader {
  const char* p;
  const char* e;
  void ws() {
    while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) ++p;
  }
  Node value(int depth = 0) {
    if (depth > 64) throw ParseError("json: nesting too deep");
    ws();
    if (p >= e) throw ParseError("json: truncated");
    Node n;
    if (*p == '{') {
      ++p;
      n.kind = Node::MAP;
      ws();
      if (p < e && *p == '}') { ++p; return n; }
      for (;;) {
        ws();
        Node k = value(depth + 1);
        ws();
        if (p >= e || *p != ':') throw ParseError("json: expected ':'");
        ++p;
        n.map.emplace_back(k.str, value(depth + 1));
        ws();
        if (p < e && *p == ',') { ++p; continue; }
        if (p < e && *p == '}') { ++p; break; }
        throw ParseError("json: expected ',' or '}'");
      }
    } else if (*p == '[') {
      ++p;
      n.kind = Node::ARR;
      ws();
      if (p < e && *p == ']') { ++p; return n; }
      for (;;) {
        n.arr.push_back(value(depth + 1));
        ws();
        if (p < e && *p == ',') { ++p; continue; }
        if (p < e && *p == ']') { ++p; break; }
        throw ParseError("json: expected ',' or ']'");
      }
    } else if (*p == '"') {
      ++p;
      n.kind = Node::STR;
      while (p < e && *p != '"') {
        if (*p == '\\' && p + 1 < e) ++p;
        n.str.push_back(*p++);
      }
      if (p >= e) throw ParseError("json: unterminated string");
      ++p;
    } else if (!std::strncmp(p, "true", 4)) { p += 4; n.kind = Node::BOOL; n.num = 1; }
    else if (!std::strncmp(p, "false", 5)) { p += 5; n.kind = Node::BOOL; n.num = 0; }
    else if (!std::strncmp(p, "null", 4)) { p += 4; n.kind = Node::NIL; }
    else {
      char* end = nullptr;
      n.num = std::strtod(p, &end);
      if (end == p) throw ParseError("json: bad value");
      n.kind = Node::NUM;
      p = end;
    }
    return n;
  }
};

// ---- YAML (the block / flow subset dynobench writes) -----------------------------
struct YamlLine {
  int indent;
  std::string text;  // without indentation and comments
};

Node yaml_scalar(const std::string& s);

Node yaml_flow(const std::string& s, size_t& i) {
  // flow sequence of scalars / nested sequences: [a, b, [c, d]]
  Node n;
  n.kind = Node::ARR;
  ++i;  // '['
  for (;;) {
    while (i < s.size() && (s[i] == ' ' || s[i] == ',')) ++i;
    if (i >= s.size()) throw ParseError("yaml: unterminated flow sequence");
    if (s[i] == ']') { ++i; break; }
    if (s[i] == '[') { n.arr.push_back(yaml_flow(s, i)); continue; }
    size_t j = i;
    while (j < s.size() && s[j] != ',' && s[j] != ']') ++j;
    std::string tok = s.substr(i, j - i);
    while (!tok.empty() && tok.back() == ' ') tok.pop_back();
    n.arr.push_back(yaml_scalar(tok));
    i = j;
  }
  return n;
}

Node yaml_scalar(const std::string& s0) {
  std::string s = s0;
  while (!s.empty() && (s.back() == ' ' || s.back() == '\r')) s.pop_back();
  Node n;
  if (s.empty() || s == "~" || s == "null") return n;
  if (s[0] == '[') {
    size_t i = 0;
    return yaml_flow(s, i);
  }
  if (s.size() >= 2 && (s[0] == '"' || s[0] == '\'') && s.back() == s[0]) {
    n.kind = Node::STR;
    n.str = s.substr(1, s.size() - 2);
    return n;
  }
  char* end = nullptr;
  double v = std::strtod(s.c_str(), &end);
  if (end && *end == '\0') {
    n.kind = Node::NUM;
    n.num = v;
    return n;
  }
  n.kind = Node::STR;
  n.str = s;
  return n;
}

struct YamlReader {
  std::vector<YamlLine> L;
  size_t i = 0;

  explicit YamlReader(const std::string& text) {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
      // strip comments outside quotes
      bool q = false;
      for (size_t k = 0; k < line.size(); ++k) {
        if (line[k] == '"') q = !q;
        if (!q && line[k] == '#' && (k == 0 || line[k - 1] == ' ')) {
          line.resize(k);
          break;
        }
      }
      size_t ind = line.find_first_not_of(' ');
      if (ind == std::string::npos) continue;
      std::string t = line.substr(ind);
      while (!t.empty() && (t.back() == ' ' || t.back() == '\r')) t.pop_back();
      if (t.empty() || t == "---") continue;
      L.push_back({(int)ind, t});
    }
  }

  // Parse the block starting at line i with indentation `ind`.
  Node block(int ind, int depth = 0) {
    if (depth > 64) throw ParseError("yaml: nesting too deep");
    Node n;
    if (i >= L.size()) return n;
    if (L[i].text[0] == '-' && (L[i].text.size() == 1 || L[i].text[1] == ' ')) {
      n.kind = Node::ARR;
      while (i < L.size() && L[i].indent == ind && L[i].text[0] == '-' &&
             (L[i].text.size() == 1 || L[i].text[1] == ' ')) {
        std::string rest = L[i].text.size() > 1 ? L[i].text.substr(2) : "";
        size_t f = rest.find_first_not_of(' ');
        rest = f == std::string::npos ? "" : rest.substr(f);
        if (rest.empty()) {
          ++i;
          if (i < L.size() && L[i].indent > ind) n.arr.push_back(block(L[i].indent, depth + 1));
          else n.arr.emplace_back();
        } else if (rest.find(": ") != std::string::npos || rest.back() == ':') {
          // "- key: value" starts an inline map at indentation ind + 2
          L[i].indent = ind + 2;
          L[i].text = rest;
          n.arr.push_back(block(ind + 2, depth + 1));
        } else {
          n.arr.push_back(yaml_scalar(rest));
          ++i;
        }
      }
      return n;
    }
    n.kind = Node::MAP;
    while (i < L.size() && L[i].indent == ind) {
      const std::string& t = L[i].text;
      size_t c = t.find(':');
      if (c == std::string::npos) throw ParseError("yaml: expected 'key: value'");
      std::string key = t.substr(0, c);
      std::string rest = t.substr(c + 1);
      size_t f = rest.find_first_not_of(' ');
      rest = f == std::string::npos ? "" : rest.substr(f);
      ++i;
      if (!rest.empty()) {
        n.map.emplace_back(key, yaml_scalar(rest));
      } else if (i < L.size() && (L[i].indent > ind || (L[i].indent == ind && L[i].text[0] == '-'))) {
        n.map.emplace_back(key, block(L[i].indent, depth + 1));
      } else {
        n.map.emplace_back(key, Node());
      }
    }
    return n;
  }
};

// ---- extraction ----------------------------------------------------------------------
std::vector<double> as_vec(const Node& n) {
  std::vector<double> v;
  if (n.kind != Node::ARR) throw ParseError("expected a vector");
  v.reserve(n.arr.size());
  for (auto& x : n.arr) {
    if (x.kind != Node::NUM && x.kind != Node::BOOL) throw ParseError("expected a number");
    v.push_back(x.num);
  }
  return v;
}

std::vector<Traj> trajs_from_doc(const Node& doc) {
  const Node* list = &doc;
  if (doc.kind == Node::MAP) {
    list = doc.get("data");
    if (!list) throw ParseError("no 'data' list");
  }
  if (list->kind != Node::ARR) throw ParseError("expected a list of trajectories");
  std::vector<Traj> out;
  out.reserve(list->arr.size());
  for (auto& t : list->arr) {
    if (t.kind != Node::MAP) throw ParseError("trajectory is not a map");
    Traj tr;
    const Node* s = t.get("states");
    if (!s || s->kind != Node::ARR) throw ParseError("trajectory without states");
    for (auto& x : s->arr) tr.states.push_back(as_vec(x));
    const Node* a = t.get("actions");
    if (a && a->kind == Node::ARR)
      for (auto& x : a->arr) tr.actions.push_back(as_vec(x));
    const Node* c = t.get("cost");
    if (c && c->kind == Node::NUM) tr.cost = c->num;
    out.push_back(std::move(tr));
  }
  return out;
}

// ---- Boost binary archive ------------------------------------------------------------
struct BoostReader {
  const uint8_t* b;
  size_t n, pos = 0;
  void need(size_t k) {
    if (n - pos < k) throw ParseError("boost archive: truncated");
  }
  uint64_t u64() { need(8); uint64_t v; std::memcpy(&v, b + pos, 8); pos += 8; return v; }
  uint32_t u32() { need(4); uint32_t v; std::memcpy(&v, b + pos, 4); pos += 4; return v; }
  double f64() { need(8); double v; std::memcpy(&v, b + pos, 8); pos += 8; return v; }
  void skip(size_t k) { need(k); pos += k; }
  std::vector<double> vec() {
    uint64_t d = u64();
    if (d > (1u << 20)) throw ParseError("boost archive: implausible vector size");
    need(8 * d);
    std::vector<double> v(d);
    if (d) std::memcpy(v.data(), b + pos, 8 * d);
    pos += 8 * d;
    return v;
  }
  std::vector<std::vector<double>> vecs(bool first) {
    uint64_t cnt = u64();
    if (cnt > (1u << 26)) throw ParseError("boost archive: implausible count");
    (void)u32();           // item version
    if (first) skip(5);    // class record of Eigen::VectorXd
    std::vector<std::vector<double>> v;
    v.reserve(cnt);
    for (uint64_t k = 0; k < cnt; ++k) v.push_back(vec());
    return v;
  }
  std::vector<Traj> read() {
    uint64_t slen = u64();
    need(slen);
    if (std::string((const char*)b + pos, slen) != "serialization::archive")
      throw ParseError("boost archive: bad signature");
    pos += slen;
    skip(2 + 5);  // library version, sizeof(int,long,float,double), endianness
    skip(13);     // class records of the trajectory collection
    const uint64_t ntraj = u64();
    if (ntraj > (1u << 26)) throw ParseError("boost archive: implausible trajectory count");
    (void)u32();
    std::vector<Traj> out;
    out.reserve(ntraj);
    for (uint64_t t = 0; t < ntraj; ++t) {
      if (t == 0) skip(10);  // class record of the trajectory
      Traj tr;
      tr.states = vecs(t == 0);
      tr.actions = vecs(false);
      tr.cost = f64();
      skip(1);      // feasible
      (void)vec();  // start
      (void)vec();  // goal
      skip(u64());  // info string
      out.push_back(std::move(tr));
    }
    return out;
  }
};

// ---- writers -----------------------------------------------------------------------------
void put_u64(std::string& s, uint64_t v) { s.append((const char*)&v, 8); }
void put_u32(std::string& s, uint32_t v) { s.append((const char*)&v, 4); }
void put_f64(std::string& s, double v) { s.append((const char*)&v, 8); }

std::string write_boost(const std::vector<Traj>& T) {
  std::string s;
  const std::string sig = "serialization::archive";
  put_u64(s, sig.size());
  s += sig;
  const uint8_t hdr[7] = {0x11, 0x00, 4, 8, 4, 8, 1};
  s.append((const char*)hdr, 7);
  s.append(13, '\0');
  put_u64(s, T.size());
  put_u32(s, 1);
  for (size_t t = 0; t < T.size(); ++t) {
    if (t == 0) {
      const uint8_t rec[10] = {0, 1, 0, 0, 0, 0, 0, 0, 0, 0};
      s.append((const char*)rec, 10);
    }
    auto vecs = [&](const std::vector<std::vector<double>>& V, bool first) {
      put_u64(s, V.size());
      put_u32(s, 0);
      if (first) s.append(5, '\0');
      for (auto& v : V) {
        put_u64(s, v.size());
        for (double x : v) put_f64(s, x);
      }
    };
    vecs(T[t].states, t == 0);
    vecs(T[t].actions, false);
    put_f64(s, T[t].cost);
    s.push_back(1);
    const std::vector<double> empty;
    const auto& st = T[t].states.empty() ? empty : T[t].states.front();
    const auto& gl = T[t].states.empty() ? empty : T[t].states.back();
    put_u64(s, st.size());
    for (double x : st) put_f64(s, x);
    put_u64(s, gl.size());
    for (double x : gl) put_f64(s, x);
    put_u64(s, 0);
  }
  return s;
}

void mp_be(std::string& s, uint64_t v, int n) {
  for (int i = n - 1; i >= 0; --i) s.push_back((char)((v >> (8 * i)) & 0xff));
}
void mp_arr(std::string& s, size_t n) {
  if (n < 16) s.push_back((char)(0x90 | n));
  else if (n < 65536) { s.push_back((char)0xdc); mp_be(s, n, 2); }
  else { s.push_back((char)0xdd); mp_be(s, n, 4); }
}
void mp_map(std::string& s, size_t n) {
  if (n < 16) s.push_back((char)(0x80 | n));
  else { s.push_back((char)0xde); mp_be(s, n, 2); }
}
void mp_str(std::string& s, const std::string& x) {
  if (x.size() < 32) s.push_back((char)(0xa0 | x.size()));
  else { s.push_back((char)0xd9); mp_be(s, x.size(), 1); }
  s += x;
}
void mp_f64(std::string& s, double v) {
  uint64_t u;
  std::memcpy(&u, &v, 8);
  s.push_back((char)0xcb);
  mp_be(s, u, 8);
}

std::string write_msgpack(const std::vector<Traj>& T) {
  std::string s;
  mp_map(s, 1);
  mp_str(s, "data");
  mp_arr(s, T.size());
  for (auto& t : T) {
    mp_map(s, 3);
    mp_str(s, "actions");
    mp_arr(s, t.actions.size());
    for (auto& v : t.actions) { mp_arr(s, v.size()); for (double x : v) mp_f64(s, x); }
    mp_str(s, "cost");
    mp_f64(s, t.cost);
    mp_str(s, "states");
    mp_arr(s, t.states.size());
    for (auto& v : t.states) { mp_arr(s, v.size()); for (double x : v) mp_f64(s, x); }
  }
  return s;
}

std::string fmt17(double x) {
  char buf[40];
  std::snprintf(buf, sizeof buf, "%.17g", x);
  return buf;
}

std::string write_yaml(const std::vector<Traj>& T) {
  std::string s;
  auto row = [&](const std::vector<double>& v) {
    s += "[";
    for (size_t k = 0; k < v.size(); ++k) { if (k) s += ","; s += fmt17(v[k]); }
    s += "]";
  };
  for (auto& t : T) {
    s += "-\n  cost: " + fmt17(t.cost) + "\n";
    s += "  num_states: " + std::to_string(t.states.size()) + "\n  states:\n";
    for (auto& v : t.states) { s += "    - "; row(v); s += "\n"; }
    s += "  num_actions: " + std::to_string(t.actions.size()) + "\n  actions:\n";
    for (auto& v : t.actions) { s += "    - "; row(v); s += "\n"; }
  }
  return s;
}

bool read_file(const char* path, std::string& out) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::ostringstream ss;
  ss << f.rdbuf();
  out = ss.str();
  return true;
}

int format_of(const char* path, int32_t format) {
  if (format != DP_FMT_AUTO) return format;
  std::string p(path);
  auto ends = [&](const char* e) {
    size_t n = std::strlen(e);
    return p.size() >= n && p.compare(p.size() - n, n, e) == 0;
  };
  if (ends(".yaml") || ends(".yml")) return DP_FMT_YAML;
  if (ends(".json")) return DP_FMT_JSON;
  if (ends(".msgpack")) return DP_FMT_MSGPACK;
  if (ends(".bin")) return DP_FMT_BOOST;
  return -1;
}

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

}  // namespace

extern "C" {

int dp_load_primitives(const char* path, int32_t format, int32_t system, int32_t max_motions, double noise,
                       uint64_t seed, dp_prim_set* out) {
  if (!path || !out) {
    dp_set_error("null argument");
    return DP_ERR_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  dp_system_desc sys;
  if (dp_system_default(system, &sys) != DP_OK) return DP_ERR_ARG;
  const int fmt = format_of(path, format);
  if (fmt < 0) {
    dp_set_error("Incompatible format for motion primitives (use .yaml .json .msgpack .bin)");
    return DP_ERR_ARG;
  }
  std::string data;
  if (!read_file(path, data)) {
    dp_set_error(std::string("cannot read ") + path);
    return DP_ERR_IO;
  }
  std::vector<Traj> T;
  try {
    if (fmt == DP_FMT_BOOST) {
      BoostReader r{(const uint8_t*)data.data(), data.size()};
      T = r.read();
    } else if (fmt == DP_FMT_MSGPACK) {
      MsgpackReader r{(const uint8_t*)data.data(), (const uint8_t*)data.data() + data.size()};
      T = trajs_from_doc(r.value());
    } else if (fmt == DP_FMT_JSON) {
      JsonReader r{data.data(), data.data() + data.size()};
      T = trajs_from_doc(r.value());
    } else {
      YamlReader r(data);
      T = trajs_from_doc(r.L.empty() ? Node() : r.block(r.L[0].indent));
    }
  } catch (const std::exception& e) {
    dp_set_error(std::string(path) + ": " + e.what());
    return DP_ERR_FORMAT;
  }
  if (T.empty()) {
    dp_set_error("no motion primitives in file");
    return DP_ERR_FORMAT;
  }
  if (max_motions > 0 && (size_t)max_motions < T.size()) T.resize(max_motions);
  const int nx = sys.nx, nu = sys.nu;
  for (size_t j = 0; j < T.size(); ++j) {
    if (T[j].states.empty()) {
      dp_set_error("primitive " + std::to_string(j) + " has no state");
      return DP_ERR_FORMAT;
    }
    for (auto& s : T[j].states)
      if ((int)s.size() != nx) {
        dp_set_error("primitive " + std::to_string(j) + ": state size does not match the system");
        return DP_ERR_FORMAT;
      }
    for (auto& a : T[j].actions)
      if ((int)a.size() != nu) {
        dp_set_error("primitive " + std::to_string(j) + ": action size does not match the system");
        return DP_ERR_FORMAT;
      }
    if (T[j].actions.size() + 1 != T[j].states.size()) {
      dp_set_error("primitive " + std::to_string(j) + ": expected one action fewer than states");
      return DP_ERR_FORMAT;
    }
  }
  if (noise != 0.0) {
    uint64_t st = seed;
    for (auto& t : T) {
      for (double& x : t.states.front()) x += noise * ((double)(splitmix64(st) >> 11) * 0x1.0p-53 * 2.0 - 1.0);
      for (double& x : t.states.back()) x += noise * ((double)(splitmix64(st) >> 11) * 0x1.0p-53 * 2.0 - 1.0);
    }
  }
  if (system == DP_SYS_QUADROTOR)
    for (auto& t : T)
      for (auto& s : t.states) {
        double n = std::sqrt(s[3] * s[3] + s[4] * s[4] + s[5] * s[5] + s[6] * s[6]);
        for (int k = 3; k < 7; ++k) s[k] /= n;
      }
  if (T.front().cost > 1e5)
    for (auto& t : T) t.cost = sys.dt * (double)t.actions.size();

  const int32_t P = (int32_t)T.size();
  int64_t S = 0;
  for (auto& t : T) S += (int64_t)t.states.size();
  out->P = P;
  out->nx = nx;
  out->nu = nu;
  out->offsets = (int32_t*)std::malloc(sizeof(int32_t) * (P + 1));
  out->states = (double*)std::malloc(sizeof(double) * S * nx);
  out->actions = (double*)std::malloc(sizeof(double) * (S - P > 0 ? S - P : 1) * nu);
  out->cost = (double*)std::malloc(sizeof(double) * P);
  if (!out->offsets || !out->states || !out->actions || !out->cost) {
    dp_free_primitives(out);
    dp_set_error("out of host memory");
    return DP_ERR_NOMEM;
  }
  int64_t s = 0, a = 0;
  for (int32_t j = 0; j < P; ++j) {
    out->offsets[j] = (int32_t)s;
    for (auto& v : T[j].states) std::memcpy(out->states + (s++) * nx, v.data(), sizeof(double) * nx);
    for (auto& v : T[j].actions) std::memcpy(out->actions + (a++) * nu, v.data(), sizeof(double) * nu);
    out->cost[j] = T[j].cost;
  }
  out->offsets[P] = (int32_t)s;
  return DP_OK;
}

// reverse_motions (dbrrt.cpp:252-339) for the translation-invariant
// unicycles: translate by minus the last position (robot.offset of the last
// state, transform_primitive(-offset)), reverse states and actions, perturb
// the new first and last states.  The reversed primitive's cost is ref_dt per
// action (traj_to_motion copies the trajectory's own cost, :3187, which the
// translated copy does not carry over).
int dp_reverse_primitives(const dp_prims* in, int32_t system, double noise, uint64_t seed, dp_prim_set* out) {
  if (!in || !out || !in->offsets || !in->states || in->P <= 0) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  if (system != DP_SYS_UNICYCLE1 && system != DP_SYS_UNICYCLE2) {
    dp_set_error("reverse of motion primitives only working for translation invariant systems (unicycles)");
    return DP_ERR_ARG;
  }
  dp_system_desc sys;
  if (dp_system_default(system, &sys) != DP_OK) return DP_ERR_ARG;
  const int nx = sys.nx, nu = sys.nu;
  if (in->nx != nx || (in->actions && in->nu != nu)) {
    dp_set_error("primitive sizes do not match the system");
    return DP_ERR_ARG;
  }
  const int32_t P = in->P;
  const int64_t S = in->offsets[P];
  out->P = P;
  out->nx = nx;
  out->nu = nu;
  out->offsets = (int32_t*)std::malloc(sizeof(int32_t) * (P + 1));
  out->states = (double*)std::malloc(sizeof(double) * S * nx);
  out->actions = (double*)std::malloc(sizeof(double) * (S - P > 0 ? S - P : 1) * nu);
  out->cost = (double*)std::malloc(sizeof(double) * P);
  if (!out->offsets || !out->states || !out->actions || !out->cost) {
    dp_free_primitives(out);
    dp_set_error("out of host memory");
    return DP_ERR_NOMEM;
  }
  uint64_t st = seed;
  for (int32_t j = 0; j <= P; ++j) out->offsets[j] = in->offsets[j];
  for (int32_t j = 0; j < P; ++j) {
    const int32_t s0 = in->offsets[j], L1 = in->offsets[j + 1] - s0;
    if (L1 < 1) {
      dp_free_primitives(out);
      dp_set_error("primitive with no state");
      return DP_ERR_ARG;
    }
    const double* last = in->states + (int64_t)(s0 + L1 - 1) * nx;
    const double ox = last[0], oy = last[1];
    for (int32_t k = 0; k < L1; ++k) {
      const double* src = in->states + (int64_t)(s0 + L1 - 1 - k) * nx;
      double* dst = out->states + (int64_t)(s0 + k) * nx;
      for (int c = 0; c < nx; ++c) dst[c] = src[c];
      dst[0] = src[0] + -ox;  // transform_primitive(-offset)
      dst[1] = src[1] + -oy;
    }
    const int32_t L = L1 - 1;
    for (int32_t k = 0; k < L; ++k)
      for (int u = 0; u < nu; ++u)
        out->actions[(int64_t)(s0 - j + k) * nu + u] =
            in->actions ? in->actions[(int64_t)(s0 - j + L - 1 - k) * nu + u] : 0.0;
    if (noise != 0.0) {
      double* f = out->states + (int64_t)s0 * nx;
      double* b = out->states + (int64_t)(s0 + L1 - 1) * nx;
      for (int c = 0; c < nx; ++c) f[c] += noise * ((double)(splitmix64(st) >> 11) * 0x1.0p-53 * 2.0 - 1.0);
      for (int c = 0; c < nx; ++c) b[c] += noise * ((double)(splitmix64(st) >> 11) * 0x1.0p-53 * 2.0 - 1.0);
    }
    out->cost[j] = sys.dt * (double)L;
  }
  return DP_OK;
}

void dp_free_primitives(dp_prim_set* set) {
  if (!set) return;
  std::free(set->offsets);
  std::free(set->states);
  std::free(set->actions);
  std::free(set->cost);
  std::memset(set, 0, sizeof(*set));
}

int dp_save_primitives(const char* path, int32_t format, const dp_prim_set* set) {
  if (!path || !set || !set->offsets || !set->states || set->P <= 0) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  const int fmt = format_of(path, format);
  std::vector<Traj> T(set->P);
  for (int32_t j = 0; j < set->P; ++j) {
    for (int32_t s = set->offsets[j]; s < set->offsets[j + 1]; ++s)
      T[j].states.emplace_back(set->states + (int64_t)s * set->nx, set->states + (int64_t)(s + 1) * set->nx);
    if (set->actions)
      for (int32_t s = set->offsets[j] - j; s < set->offsets[j + 1] - j - 1; ++s)
        T[j].actions.emplace_back(set->actions + (int64_t)s * set->nu, set->actions + (int64_t)(s + 1) * set->nu);
    T[j].cost = set->cost ? set->cost[j] : 1e8;
  }
  std::string data;
  if (fmt == DP_FMT_BOOST) data = write_boost(T);
  else if (fmt == DP_FMT_MSGPACK) data = write_msgpack(T);
  else if (fmt == DP_FMT_YAML) data = write_yaml(T);
  else {
    dp_set_error("writing supports .bin (Boost), .msgpack and .yaml");
    return DP_ERR_ARG;
  }
  std::ofstream f(path, std::ios::binary);
  if (!f || !f.write(data.data(), (std::streamsize)data.size())) {
    dp_set_error(std::string("cannot write ") + path);
    return DP_ERR_IO;
  }
  return DP_OK;
}

// ---- problem files ------------------------------------------------------------------
// dynobench's problem YAML (the layout benchmark.py's create_empty_env
// writes, benchmark/benchmark.py:330-350):
//   environment: {min: [..], max: [..], obstacles: [{type, center, size}]}
//   robots: [{type: <model name>, start: [..], goal: [..]}]
// Box sizes are full side lengths; a sphere's size is [radius].

static int system_of_model(const std::string& t) {
  auto starts = [&](const char* p) { return t.compare(0, std::strlen(p), p) == 0; };
  if (starts("unicycle1")) return DP_SYS_UNICYCLE1;
  if (starts("unicycle2")) return DP_SYS_UNICYCLE2;
  if (starts("quadrotor") || starts("quad3d")) return DP_SYS_QUADROTOR;
  if (starts("car_first_order_with_1_trailers") || starts("car1")) return DP_SYS_CAR_TRAILER;
  return -1;
}

int dp_load_problem(const char* path, dp_problem* out) {
  if (!path || !out) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  std::memset(out, 0, sizeof(*out));
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    dp_set_error(std::string("cannot open ") + path);
    return DP_ERR_IO;
  }
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string data = ss.str();
  std::vector<double> boxes, spheres;
  try {
    YamlReader r(data);
    if (r.L.empty()) throw ParseError("empty problem file");
    Node doc = r.block(r.L[0].indent);
    const Node* env = doc.get("environment");
    const Node* robots = doc.get("robots");
    if (!env || !robots || robots->kind != Node::ARR || robots->arr.empty())
      throw ParseError("expected 'environment' and a non-empty 'robots' list");
    const Node& rb = robots->arr[0];
    const Node* type = rb.get("type");
    if (!type || type->kind != Node::STR) throw ParseError("robots[0].type missing");
    const int sys = system_of_model(type->str);
    if (sys < 0) throw ParseError("unsupported robot type " + type->str);
    dp_system_desc d;
    dp_system_default(sys, &d);
    out->system = sys;
    out->nx = d.nx;
    std::snprintf(out->robot_type, sizeof(out->robot_type), "%s", type->str.c_str());
    for (const char* key : {"start", "goal"}) {
      const Node* v = rb.get(key);
      if (!v) throw ParseError(std::string("robots[0].") + key + " missing");
      std::vector<double> x = as_vec(*v);
      if ((int)x.size() != d.nx) throw ParseError(std::string(key) + " has the wrong dimension");
      std::memcpy(key[0] == 's' ? out->start : out->goal, x.data(), sizeof(double) * x.size());
    }
    const Node* mn = env->get("min");
    const Node* mx = env->get("max");
    if (!mn || !mx) throw ParseError("environment.min / max missing");
    std::vector<double> lo = as_vec(*mn), hi = as_vec(*mx);
    if (lo.size() != hi.size() || lo.size() < 2 || lo.size() > 3) throw ParseError("environment bounds");
    for (size_t k = 0; k < lo.size(); ++k) {
      out->lo[k] = lo[k];
      out->hi[k] = hi[k];
    }
    const Node* obs = env->get("obstacles");
    if (obs && obs->kind == Node::ARR)
      for (const Node& o : obs->arr) {
        const Node* ot = o.get("type");
        const Node* c = o.get("center");
        const Node* sz = o.get("size");
        if (!ot || !c || !sz) throw ParseError("obstacle needs type, center, size");
        std::vector<double> cv = as_vec(*c), sv = as_vec(*sz);
        cv.resize(3, 0.0);
        if (ot->str == "box") {
          sv.resize(3, 0.0);
          for (int k = 0; k < 3; ++k) boxes.push_back(cv[k]);
          for (int k = 0; k < 3; ++k) boxes.push_back(0.5 * sv[k]);
        } else if (ot->str == "sphere") {
          if (sv.empty()) throw ParseError("sphere size");
          for (int k = 0; k < 3; ++k) spheres.push_back(cv[k]);
          spheres.push_back(sv[0]);
        } else {
          throw ParseError("unsupported obstacle type " + ot->str);
        }
      }
  } catch (const std::exception& e) {
    dp_set_error(std::string(path) + ": " + e.what());
    return DP_ERR_FORMAT;
  }
  out->n_boxes = (int32_t)(boxes.size() / 6);
  out->n_spheres = (int32_t)(spheres.size() / 4);
  out->boxes = (double*)std::malloc(sizeof(double) * (boxes.size() + 1));
  out->spheres = (double*)std::malloc(sizeof(double) * (spheres.size() + 1));
  if (!out->boxes || !out->spheres) {
    dp_free_problem(out);
    dp_set_error("out of host memory");
    return DP_ERR_NOMEM;
  }
  if (!boxes.empty()) std::memcpy(out->boxes, boxes.data(), sizeof(double) * boxes.size());
  if (!spheres.empty()) std::memcpy(out->spheres, spheres.data(), sizeof(double) * spheres.size());
  return DP_OK;
}

void dp_free_problem(dp_problem* p) {
  if (!p) return;
  std::free(p->boxes);
  std::free(p->spheres);
  p->boxes = nullptr;
  p->spheres = nullptr;
  p->n_boxes = p->n_spheres = 0;
}

int dp_save_problem(const char* path, const dp_problem* p) {
  if (!path || !p || p->nx <= 0 || p->nx > DP_MAX_NX || (p->n_boxes && !p->boxes) || (p->n_spheres && !p->spheres)) {
    dp_set_error("invalid argument");
    return DP_ERR_ARG;
  }
  const int dims = p->system == DP_SYS_QUADROTOR ? 3 : 2;
  std::ostringstream o;
  o.precision(17);
  auto vec = [&](const double* v, int n) {
    o << "[";
    for (int k = 0; k < n; ++k) o << (k ? ", " : "") << v[k];
    o << "]";
  };
  o << "environment:\n  min: ";
  vec(p->lo, dims);
  o << "\n  max: ";
  vec(p->hi, dims);
  o << "\n  obstacles:\n";
  for (int32_t i = 0; i < p->n_boxes; ++i) {
    const double* b = p->boxes + 6 * i;
    double size[3] = {2.0 * b[3], 2.0 * b[4], 2.0 * b[5]};
    o << "    - type: box\n      center: ";
    vec(b, dims);
    o << "\n      size: ";
    vec(size, dims);
    o << "\n";
  }
  for (int32_t i = 0; i < p->n_spheres; ++i) {
    const double* s = p->spheres + 4 * i;
    o << "    - type: sphere\n      center: ";
    vec(s, 3);
    o << "\n      size: ";
    vec(s + 3, 1);
    o << "\n";
  }
  o << "robots:\n  - type: " << (p->robot_type[0] ? p->robot_type : "unicycle1_v0") << "\n    start: ";
  vec(p->start, p->nx);
  o << "\n    goal: ";
  vec(p->goal, p->nx);
  o << "\n";
  std::ofstream f(path, std::ios::binary);
  const std::string data = o.str();
  if (!f || !f.write(data.data(), (std::streamsize)data.size())) {
    dp_set_error(std::string("cannot write ") + path);
    return DP_ERR_IO;
  }
  return DP_OK;
}

}  // extern "C"
