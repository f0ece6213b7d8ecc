// config.cpp -- data/default.cfg and plan-JSON loaders with the reference's
// semantics (src/io/io_json.cpp), on a small JSON reader of our own.
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>

#include "host.hpp"
#include "solver_fim.hpp"

namespace fim {
namespace {

// ---------------------------------------------------------------------------
// JSON value + recursive-descent reader.  Objects keep keys sorted (std::map),
// the iteration order the reference's JSON library uses.
// ---------------------------------------------------------------------------
struct JVal {
  enum Kind { Null, Bool, Num, Str, Arr, Obj } k = Null;
  bool b = false;
  double n = 0.0;
  std::string s;
  std::vector<JVal> a;
  std::map<std::string, JVal> o;

  bool has(const std::string &key) const { return k == Obj && o.count(key); }
  const JVal &at(const std::string &key) const {
    auto it = o.find(key);
    if (k != Obj || it == o.end()) throw Error(FIM_EINVAL, "plan: missing key '" + key + "'");
    return it->second;
  }
  double num() const {
    if (k == Num) return n;
    if (k == Bool) return b ? 1.0 : 0.0;
    throw Error(FIM_EINVAL, "plan: expected a number");
  }
  const std::string &str() const {
    if (k != Str) throw Error(FIM_EINVAL, "plan: expected a string");
    return s;
  }
};

struct JReader {
  const std::string &t;
  size_t i = 0;
  explicit JReader(const std::string &text) : t(text) {}

  [[noreturn]] void fail(const char *what) {
    throw Error(FIM_EIO, std::string("plan: JSON parse error (") + what + ") at offset " +
                             std::to_string(i));
  }
  void ws() {
    while (i < t.size() && (t[i] == ' ' || t[i] == '\n' || t[i] == '\r' || t[i] == '\t')) ++i;
  }
  bool lit(const char *w) {
    size_t n = 0;
    while (w[n]) ++n;
    if (t.compare(i, n, w) == 0) {
      i += n;
      return true;
    }
    return false;
  }
  std::string string() {
    if (t[i] != '"') fail("expected string");
    ++i;
    std::string out;
    while (i < t.size() && t[i] != '"') {
      char c = t[i++];
      if (c == '\\') {
        if (i >= t.size()) fail("bad escape");
        const char e = t[i++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (i + 4 > t.size()) fail("bad \\u escape");
            const unsigned cp = (unsigned)std::strtoul(t.substr(i, 4).c_str(), nullptr, 16);
            i += 4;
            if (cp < 0x80) {
              out += (char)cp;
            } else if (cp < 0x800) {
              out += (char)(0xC0 | (cp >> 6));
              out += (char)(0x80 | (cp & 0x3F));
            } else {
              out += (char)(0xE0 | (cp >> 12));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (i >= t.size()) fail("unterminated string");
    ++i;
    return out;
  }
  JVal value(int depth = 0) {
    if (depth > 256) fail("nesting too deep");
    ws();
    if (i >= t.size()) fail("unexpected end");
    JVal v;
    const char c = t[i];
    if (c == '{') {
      ++i;
      v.k = JVal::Obj;
      ws();
      if (t[i] == '}') {
        ++i;
        return v;
      }
      for (;;) {
        ws();
        std::string key = string();
        ws();
        if (t[i] != ':') fail("expected ':'");
        ++i;
        v.o[key] = value(depth + 1);
        ws();
        if (t[i] == ',') {
          ++i;
          continue;
        }
        if (t[i] == '}') {
          ++i;
          break;
        }
        fail("expected ',' or '}'");
      }
    } else if (c == '[') {
      ++i;
      v.k = JVal::Arr;
      ws();
      if (t[i] == ']') {
        ++i;
        return v;
      }
      for (;;) {
        v.a.push_back(value(depth + 1));
        ws();
        if (t[i] == ',') {
          ++i;
          continue;
        }
        if (t[i] == ']') {
          ++i;
          break;
        }
        fail("expected ',' or ']'");
      }
    } else if (c == '"') {
      v.k = JVal::Str;
      v.s = string();
    } else if (lit("true")) {
      v.k = JVal::Bool;
      v.b = true;
    } else if (lit("false")) {
      v.k = JVal::Bool;
      v.b = false;
    } else if (lit("null")) {
      v.k = JVal::Null;
    } else {
      const char *p = t.c_str() + i;
      char *end = nullptr;
      errno = 0;
      v.n = std::strtod(p, &end);
      if (end == p) fail("bad token");
      i += (size_t)(end - p);
      v.k = JVal::Num;
    }
    return v;
  }
};

std::string read_file(const std::string &path, const char *what) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw Error(FIM_EIO, std::string(what) + ": file not found: " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

// non-empty lines (src/io/io_base.cpp:7-31 drops empty ones)
std::vector<std::string> lines_of(const std::string &text) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : text) {
    if (c == '\n') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

}  // namespace

Options load_options(const std::string &cfg_path) {
  const std::vector<std::string> cfg = lines_of(read_file(cfg_path, "load_options"));
  if (cfg.size() < 9) throw Error(FIM_EINVAL, "load_options: expected 9 non-empty lines in " + cfg_path);
  Options o;
  if (cfg[0] == "false")
    o.verbose = 0;
  else if (cfg[0] == "true")
    o.verbose = 1;
  else
    throw Error(FIM_EINVAL, "invalid value for verbose");
  o.method_id = 1;  // line 2 is read but the solver is always FIM
  if (cfg[2] == "lin")
    o.solvetype = 0;
  else if (cfg[2] == "vol")
    o.solvetype = 1;
  else
    throw Error(FIM_EINVAL, "invalid value for slvtype");
  if (cfg[4] == "false")
    o.rent_en = 0;
  else if (cfg[4] == "true")
    o.rent_en = 1;
  else
    throw Error(FIM_EINVAL, "invalid value for re_flag");
  char *end = nullptr;
  o.dt = std::strtod(cfg[7].c_str(), &end) * 1000;
  if (*end) throw Error(FIM_EINVAL, "invalid value for dt");
  o.n_steps = std::strtol(cfg[8].c_str(), &end, 10);
  if (*end) throw Error(FIM_EINVAL, "invalid value for t_max");
  return o;
}

Plan load_plan(const std::string &plan_path, Options &opts) {
  const std::string text = read_file(plan_path, "read_jdata");
  JReader rd(text);
  const JVal root = rd.value();
  if (root.k != JVal::Obj) throw Error(FIM_EIO, "plan: top level must be an object");
  Plan plan;
  // config: {name: {func, tags}} -> func -> tags (src/io/io_json.cpp:17-32)
  std::map<std::string, std::vector<int32_t>> configs;
  for (const auto &kv : root.at("config").o)
    for (const auto &t : kv.second.at("tags").a) configs[kv.second.at("func").str()].push_back((int32_t)t.num());
  // functions: CV.CVref and EP.apd per active func (src/io/io_json.cpp:82-128)
  const JVal &funcs = root.at("functions");
  for (const auto &cf : configs) {
    const JVal &fnc = funcs.at(cf.first);
    const bool active = fnc.has("EP") && ((fnc.at("EP").k == JVal::Obj && !fnc.at("EP").o.empty()) ||
                                          (fnc.at("EP").k == JVal::Arr && !fnc.at("EP").a.empty()));
    if (!active) continue;
    Region r;
    r.func = cf.first;
    r.tags = cf.second;
    const JVal &ref = fnc.at("CV").at("CVref");
    r.vf = ref.at("vf").num();
    r.vs = ref.at("vs").num();
    r.vn = ref.at("vn").num();
    r.apd = fnc.at("EP").has("apd") ? fnc.at("EP").at("apd").num() * 1000.0 : 3e5;
    plan.regions.push_back(r);
  }
  // simulation (src/io/io_json.cpp:130-223)
  const JVal &sim = root.at("simulation");
  opts.dt = sim.at("deltatime").num() * 1000.0;
  opts.n_steps = (int64_t)sim.at("timesteps").num();
  if (!sim.has("stimuli")) return plan;  // reference: warning, no events
  for (const auto &kv : sim.at("stimuli").o) {
    const JVal &st = kv.second;
    const int active = (int)st.at("active").num();
    const std::string type = st.at("type").str();
    if (!active) continue;
    if (type != "Transmembrane Current" && type != "Transmembrane Voltage clamp")
      throw Error(FIM_EINVAL, "read_sim_from_json: unknown stim type: " + type);
    Event ev;
    ev.active = active;
    const JVal &pr = st.at("protocol");
    ev.tstart = pr.at("start").num() * 1000.0;
    ev.dur = pr.at("duration").num() * 1000.0;
    const double bcl = pr.at("bcl").num() * 1000.0;
    const int npls = (int)pr.at("npls").num();
    const JVal &el = st.at("electrode");
    if (el.at("type").str() == "vtxdata")
      for (const auto &v : el.at("vtxdata").a) ev.vtx.push_back((int64_t)v.num());
    plan.events.push_back(ev);
    for (int k = 1; k < npls; ++k) {
      Event e2 = ev;
      e2.tstart = ev.tstart + k * bcl;
      plan.events.push_back(e2);
    }
  }
  return plan;
}

}  // namespace fim
