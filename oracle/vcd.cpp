// CPU restatement of GTPower's VCD reading and per-bit event store build.
//
// TEST INFRASTRUCTURE ONLY: the oracle for the device VCD ingest and the CPU
// baseline of its benchmark.  It follows, token by token:
//   * VcdReader::read / getToken / parseTimescale / parseVar / parseScope /
//     parseUpscope / parseVarValues (power/VcdReader.cc:80-131, :145-310,
//     :312-359), single-threaded over the text as the reference reads its gz
//     stream;
//   * Vcd::makeVar / varAppendValue / varAppendBusValue (power/Vcd.cc:165-232):
//     values are kept per VCD id, a change equal to the id's last kept value is
//     dropped, and an id whose first change is after time 0 starts with X (a
//     scalar) or 0 (a bus) at time 0;
//   * VcdValue::value(bit) (power/Vcd.cc:271-277) and the value table of the
//     CUDA event store (power/cuda/CudaPower.cu:1129-1145): one event stream per
//     (id, bit), bit b of a bus value being (value >> b) & 1.
// Deliberate choices, shared with the device ingest: 'U' maps to 2 like X (the
// reference's table leaves it unmapped); a stream's times must increase
// strictly and stay below 2^29 - 1 ticks (the reference only asserts the first
// in debug builds); every condition the reference reports as a file error is
// an error here too.
#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

struct VcdVal {
  int64_t time;
  char value;       // 0 1 X Z U (upper case) or '\0' for a bus value
  uint64_t bus;
};

struct Var {
  std::string name, id;
  int width;
  int64_t uid;
};

struct Vcd {
  std::vector<Var> vars;
  std::unordered_map<std::string, int64_t> uid_of;   // id -> unique id index
  std::vector<int> uid_width;                         // max declared width
  std::vector<std::vector<VcdVal>> values;            // per unique id
  double time_scale = 1.0, unit_scale = 1.0;
  int64_t time_max = 0;
  std::string error;
  // per-bit streams
  std::vector<int64_t> stream_base, offsets, times;
  std::vector<uint8_t> vals;
};

struct Reader {
  const char* p;
  const char* end;
  Vcd* v;
  int64_t time = 0;
  std::vector<std::string> scope;

  int getc_() { return p < end ? (unsigned char)*p++ : EOF; }

  // power/VcdReader.cc:287-310: skip whitespace, read up to whitespace; the
  // character after the token is consumed; EOF right after a token drops it
  std::string token() {
    std::string t;
    int ch = getc_();
    while (ch != EOF && isspace(ch)) ch = getc_();
    while (ch != EOF && !isspace(ch)) {
      t.push_back((char)ch);
      ch = getc_();
    }
    if (ch == EOF) return "";
    return t;
  }

  std::vector<std::string> stmt_tokens() {
    std::vector<std::string> ts;
    std::string t = token();
    while (!t.empty() && t != "$end") {
      ts.push_back(t);
      t = token();
    }
    return ts;
  }

  bool fail(const std::string& m) {
    if (v->error.empty()) v->error = m;
    return false;
  }

  bool set_unit(const std::string& u) {
    if (u == "fs") v->unit_scale = 1e-15;
    else if (u == "ps") v->unit_scale = 1e-12;
    else if (u == "ns") v->unit_scale = 1e-9;
    else return fail("unknown timescale unit " + u);
    return true;
  }

  bool timescale() {
    std::vector<std::string> ts = stmt_tokens();
    try {
      if (ts.size() == 1) {
        size_t last = 0;
        const double s = std::stod(ts[0], &last);
        if (!set_unit(ts[0].substr(last))) return false;
        v->time_scale = s * v->unit_scale;
      } else if (ts.size() == 2) {
        if (!set_unit(ts[1])) return false;
        v->time_scale = std::stod(ts[0]) * v->unit_scale;
      } else {
        return fail("timescale syntax error");
      }
    } catch (...) {
      return fail("timescale syntax error");
    }
    return true;
  }

  bool var() {
    static const char* types[] = {"wire", "reg", "parameter", "integer", "real", "supply0", "supply1", "time",
                                  "tri", "triand", "trior", "trireg", "tri0", "tri1", "wand", "wor"};
    std::vector<std::string> ts = stmt_tokens();
    if (ts.size() != 4 && ts.size() != 5) return fail("variable syntax error");
    bool known = false;
    for (const char* t : types) known = known || ts[0] == t;
    if (!known) return true;  // the reference warns and skips the variable
    int width = 0;
    try {
      width = std::stoi(ts[1]);
    } catch (...) {
      return fail("variable width syntax error");
    }
    if (width < 1 || width > 64) return fail("variable width outside [1, 64]");
    std::string name;
    for (const std::string& c : scope) name += c + "/";
    std::string leaf;
    for (char c : ts[3]) {
      if (c == '/') leaf += "\\/";
      else leaf += c;
    }
    name += leaf;
    if (ts.size() == 5) {
      if (name[0] == '\\') name += ' ';
      name += ts[4];
    }
    const std::string& id = ts[2];
    auto it = v->uid_of.find(id);
    int64_t uid;
    if (it == v->uid_of.end()) {
      uid = (int64_t)v->uid_width.size();
      v->uid_of[id] = uid;
      v->uid_width.push_back(width);
      v->values.emplace_back();
    } else {
      uid = it->second;
      v->uid_width[uid] = std::max(v->uid_width[uid], width);
      v->values[uid].clear();  // makeVar clears the id's values
    }
    v->vars.push_back(Var{name, id, width, uid});
    return true;
  }

  // power/Vcd.cc:203-216
  bool append(int64_t uid, char value) {
    std::vector<VcdVal>& vs = v->values[uid];
    if (vs.empty() || vs.back().value != value) {
      if (!vs.empty() && vs.back().time >= time) return fail("value changes of an id not increasing in time");
      if (vs.empty() && time != 0) vs.push_back(VcdVal{0, 'X', 0});
      vs.push_back(VcdVal{time, value, 0});
    }
    return true;
  }

  // power/Vcd.cc:218-232
  bool append_bus(int64_t uid, uint64_t bus) {
    std::vector<VcdVal>& vs = v->values[uid];
    if (vs.empty() || vs.back().bus != bus) {
      if (!vs.empty() && vs.back().time >= time) return fail("value changes of an id not increasing in time");
      if (vs.empty() && time != 0) vs.push_back(VcdVal{0, '\0', 0});
      vs.push_back(VcdVal{time, '\0', bus});
    }
    return true;
  }

  int64_t lookup(const std::string& id) {
    auto it = v->uid_of.find(id);
    return it == v->uid_of.end() ? -1 : it->second;
  }

  // power/VcdReader.cc:238-285: to the end of the text
  bool values() {
    std::string t = token();
    while (!t.empty()) {
      const char c0 = (char)toupper((unsigned char)t[0]);
      if (c0 == '#' && t.size() > 1) {
        size_t used = 0;
        try {
          time = std::stoll(t.substr(1), &used);
        } catch (...) {
          return fail("time syntax error: " + t);
        }
      } else if (c0 == '0' || c0 == '1' || c0 == 'X' || c0 == 'U' || c0 == 'Z') {
        const int64_t uid = lookup(t.substr(1));
        if (uid < 0) return fail("unknown variable " + t.substr(1));
        if (!append(uid, c0)) return false;
      } else if (c0 == 'B') {
        const char c1 = t.size() > 1 ? (char)toupper((unsigned char)t[1]) : '\0';
        const std::string id = token();
        const int64_t uid = lookup(id);
        if (uid < 0) return fail("unknown variable " + id);
        uint64_t bus = 0;
        if (!(c1 == 'X' || c1 == 'U' || c1 == 'Z')) bus = (uint64_t)std::strtol(t.c_str() + 1, nullptr, 2);
        if (!append_bus(uid, bus)) return false;
      }
      t = token();
    }
    v->time_max = time;
    return true;
  }

  bool read() {
    std::string t = token();
    while (!t.empty()) {
      if (t == "$date" || t == "$comment" || t == "$version" || t == "$enddefinitions") {
        stmt_tokens();
      } else if (t == "$timescale") {
        if (!timescale()) return false;
      } else if (t == "$var") {
        if (!var()) return false;
      } else if (t == "$scope") {
        std::vector<std::string> ts = stmt_tokens();
        if (ts.size() < 2) return fail("scope syntax error");
        scope.push_back(ts[1]);
      } else if (t == "$upscope") {
        stmt_tokens();
        if (scope.empty()) return fail("upscope without scope");
        scope.pop_back();
      } else if (t == "$dumpall" || t == "$dumpvars" || t[0] != '$') {
        return values();  // the token that starts the value changes is not one of them
      } else {
        return fail("unhandled vcd command " + t);
      }
      t = token();
    }
    return true;
  }
};

uint8_t code(char c) {
  // power/cuda/CudaPower.cu:1129-1145; 'U' like 'X'
  switch (c) {
    case '0': return 0;
    case '1': return 1;
    case 'Z': return 3;
    default: return 2;
  }
}

void build_streams(Vcd& v) {
  const int64_t U = (int64_t)v.uid_width.size();
  v.stream_base.assign(U + 1, 0);
  for (int64_t u = 0; u < U; ++u) v.stream_base[u + 1] = v.stream_base[u] + v.uid_width[u];
  const int64_t S = v.stream_base[U];
  v.offsets.assign(S + 1, 0);
  for (int64_t u = 0; u < U; ++u)
    for (int b = 0; b < v.uid_width[u]; ++b)
      v.offsets[v.stream_base[u] + b + 1] = (int64_t)v.values[u].size();
  for (int64_t s = 0; s < S; ++s) v.offsets[s + 1] += v.offsets[s];
  v.times.resize(v.offsets[S]);
  v.vals.resize(v.offsets[S]);
  for (int64_t u = 0; u < U; ++u)
    for (int b = 0; b < v.uid_width[u]; ++b) {
      int64_t o = v.offsets[v.stream_base[u] + b];
      for (const VcdVal& x : v.values[u]) {
        v.times[o] = x.time;
        // power/Vcd.cc:271-277
        v.vals[o] = x.value ? code(x.value) : (uint8_t)((x.bus >> b) & 1u);
        ++o;
      }
    }
}

}  // namespace

extern "C" {

// Parse a VCD text; nullptr on error with the message in err.
void* oracle_vcd_parse(const char* text, int64_t len, char* err, int64_t err_len) {
  Vcd* v = new Vcd();
  Reader r{text, text + len, v};
  bool ok = r.read();
  if (ok) {
    for (const std::vector<VcdVal>& vs : v->values)
      for (const VcdVal& x : vs)
        if (x.time >= (int64_t)((1u << 29) - 1) || x.time < 0) {
          ok = false;
          v->error = "value change time outside [0, 2^29 - 1)";
        }
  }
  if (!ok) {
    if (err && err_len > 0) {
      std::strncpy(err, v->error.c_str(), (size_t)err_len - 1);
      err[err_len - 1] = 0;
    }
    delete v;
    return nullptr;
  }
  build_streams(*v);
  return v;
}

// out = {n_var, n_id, n_stream, n_event, t_max}; time_scale = seconds per tick
void oracle_vcd_info(void* h, int64_t* out, double* time_scale) {
  const Vcd* v = (const Vcd*)h;
  out[0] = (int64_t)v->vars.size();
  out[1] = (int64_t)v->uid_width.size();
  out[2] = v->stream_base.back();
  out[3] = v->offsets.back();
  out[4] = v->time_max;
  *time_scale = v->time_scale;
}

const char* oracle_vcd_var_name(void* h, int64_t i) { return ((Vcd*)h)->vars[i].name.c_str(); }
const char* oracle_vcd_var_id(void* h, int64_t i) { return ((Vcd*)h)->vars[i].id.c_str(); }
int32_t oracle_vcd_var_width(void* h, int64_t i) { return ((Vcd*)h)->vars[i].width; }
int64_t oracle_vcd_var_stream(void* h, int64_t i) {
  const Vcd* v = (const Vcd*)h;
  return v->stream_base[v->vars[i].uid];
}

void oracle_vcd_events(void* h, int64_t* offsets, int64_t* times, uint8_t* values) {
  const Vcd* v = (const Vcd*)h;
  std::memcpy(offsets, v->offsets.data(), v->offsets.size() * sizeof(int64_t));
  if (!v->times.empty()) {
    std::memcpy(times, v->times.data(), v->times.size() * sizeof(int64_t));
    std::memcpy(values, v->vals.data(), v->vals.size());
  }
}

void oracle_vcd_free(void* h) { delete (Vcd*)h; }

}  // extern "C"
