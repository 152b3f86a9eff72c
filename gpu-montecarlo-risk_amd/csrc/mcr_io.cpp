// mcr_io.cpp -- the data formats around the risk path (include/mcr_io.h):
// binary inputs (src/utilities/load_binary.c, preprocessing/save_binary.py),
// the YAML configuration (src/utilities/load_config.c) and the results CSV
// (src/utilities/csv_writer.c).  Host code only; nothing here prints.
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mcr_hip.h"
#include "../../include/mcr_io.h"
#include "mcr_host.h"

using namespace mcr::host;

namespace {

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

// mkdir -p
int make_dirs(const std::string& dir) {
  if (dir.empty()) return MCR_OK;
  std::string cur;
  size_t pos = 0;
  while (pos != std::string::npos) {
    pos = dir.find('/', pos + 1);
    cur = dir.substr(0, pos);
    if (cur.empty()) continue;
    if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST) {
      set_error("cannot create directory %s: %s", cur.c_str(), strerror(errno));
      return MCR_E_ARG;
    }
  }
  return MCR_OK;
}

int write_file(const std::string& path, const void* const* parts, const size_t* sizes, int nparts) {
  File out;
  out.f = fopen(path.c_str(), "wb");
  if (!out.f) {
    set_error("cannot open %s for writing: %s", path.c_str(), strerror(errno));
    return MCR_E_ARG;
  }
  for (int i = 0; i < nparts; ++i)
    if (sizes[i] && fwrite(parts[i], 1, sizes[i], out.f) != sizes[i]) {
      set_error("short write to %s", path.c_str());
      return MCR_E_ARG;
    }
  if (fclose(out.f) != 0) {
    out.f = nullptr;
    set_error("cannot close %s: %s", path.c_str(), strerror(errno));
    return MCR_E_ARG;
  }
  out.f = nullptr;
  return MCR_OK;
}

// ---------------------------------------------------------------- YAML ---
std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r')) ++a;
  while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r')) --b;
  return s.substr(a, b - a);
}

// Drops a '#' comment: '#' at the start or after whitespace, outside quotes.
std::string strip_comment(const std::string& s) {
  char quote = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (quote) {
      if (c == quote) {
        if (quote == '\'' && i + 1 < s.size() && s[i + 1] == '\'') {
          ++i;  // '' inside single quotes
        } else {
          quote = 0;
        }
      } else if (c == '\\' && quote == '"') {
        ++i;
      }
    } else if (c == '"' || c == '\'') {
      quote = c;
    } else if (c == '#' && (i == 0 || s[i - 1] == ' ' || s[i - 1] == '\t')) {
      return s.substr(0, i);
    }
  }
  return s;
}

// Plain, "double" (\" \\ \n \t escapes) or 'single' ('' escape) scalar.
bool unquote(const std::string& v, std::string* out) {
  if (v.size() >= 2 && v.front() == '"' && v.back() == '"') {
    std::string r;
    for (size_t i = 1; i + 1 < v.size(); ++i) {
      char c = v[i];
      if (c == '\\' && i + 2 < v.size()) {
        const char e = v[++i];
        c = e == 'n' ? '\n' : e == 't' ? '\t' : e;
      }
      r.push_back(c);
    }
    *out = r;
    return true;
  }
  if (v.size() >= 2 && v.front() == '\'' && v.back() == '\'') {
    std::string r;
    for (size_t i = 1; i + 1 < v.size(); ++i) {
      r.push_back(v[i]);
      if (v[i] == '\'' && i + 2 < v.size() && v[i + 1] == '\'') ++i;
    }
    *out = r;
    return true;
  }
  if (!v.empty() && (v.front() == '"' || v.front() == '\'')) return false;  // unterminated
  *out = v;
  return true;
}

// Position of the ':' that separates a key from its value (followed by a
// space or the end), outside quotes; npos if none.
size_t key_colon(const std::string& s) {
  char quote = 0;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (quote) {
      if (c == quote) quote = 0;
    } else if (c == '"' || c == '\'') {
      quote = c;
    } else if (c == ':' && (i + 1 == s.size() || s[i + 1] == ' ')) {
      return i;
    }
  }
  return std::string::npos;
}

// Splits a flow collection body on commas outside quotes.
std::vector<std::string> split_flow(const std::string& body) {
  std::vector<std::string> out;
  std::string cur;
  char quote = 0;
  for (char c : body) {
    if (quote) {
      if (c == quote) quote = 0;
      cur.push_back(c);
    } else if (c == '"' || c == '\'') {
      quote = c;
      cur.push_back(c);
    } else if (c == ',') {
      out.push_back(trim(cur));
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  if (!trim(cur).empty()) out.push_back(trim(cur));
  return out;
}

char* dup(const std::string& s) {
  char* p = (char*)malloc(s.size() + 1);
  if (p) memcpy(p, s.c_str(), s.size() + 1);
  return p;
}

void set_string(char** field, const std::string& v) {
  free(*field);
  *field = dup(v);
}

// Flat keys as the reference assigns them (load_config.c:243-268).
void assign_scalar(mcr_config* c, const std::string& key, const std::string& v) {
  if (key == "start") {
    set_string(&c->start, v);
  } else if (key == "end") {
    set_string(&c->end, v);
  } else if (key == "x") {
    c->x = atof(v.c_str());
  } else if (key == "N") {
    c->N = atoi(v.c_str());
  } else if (key == "M") {
    c->M = atoi(v.c_str());
  } else if (key == "k") {
    c->k = atoi(v.c_str());
  } else if (key == "random_seed") {
    c->random_seed = (unsigned long)atol(v.c_str());
  } else if (key == "train_ratio") {
    c->train_ratio = atof(v.c_str());
  } else if (key == "comment") {
    set_string(&c->comment, v);
  }
}

void add_model(mcr_config* c, const std::string& m) {
  char** nm = (char**)realloc(c->models, (c->num_models + 1) * sizeof(char*));
  if (!nm) return;
  c->models = nm;
  c->models[c->num_models++] = dup(m);
}

void add_index(mcr_config* c, const std::string& name, const std::string& file) {
  mcr_index* ni = (mcr_index*)realloc(c->indices, (c->num_indices + 1) * sizeof(mcr_index));
  if (!ni) return;
  c->indices = ni;
  c->indices[c->num_indices].name = dup(name);
  c->indices[c->num_indices].file = dup(file);
  c->num_indices++;
}

void clear_list(mcr_config* c, const std::string& key) {
  if (key == "models") {
    for (int i = 0; i < c->num_models; ++i) free(c->models[i]);
    free(c->models);
    c->models = nullptr;
    c->num_models = 0;
  } else if (key == "indices") {
    for (int i = 0; i < c->num_indices; ++i) {
      free(c->indices[i].name);
      free(c->indices[i].file);
    }
    free(c->indices);
    c->indices = nullptr;
    c->num_indices = 0;
  }
}

// ----------------------------------------------------------------- CSV ---
std::mutex g_csv_mu;

std::string csv_field(const char* s) {
  if (!s || !*s) return std::string();
  bool need = false;
  for (const char* p = s; *p; ++p)
    if (*p == ',' || *p == '"' || *p == '\n' || *p == '\r') need = true;
  if (!need) return std::string(s);
  std::string r = "\"";
  for (const char* p = s; *p; ++p) {
    if (*p == '"') r.push_back('"');
    r.push_back(*p);
  }
  r.push_back('"');
  return r;
}

constexpr const char* kCsvHeader =
    "iterationnumber,timestamp,execution_time_ms,comment,start_date,end_date,train_ratio,M,k,x,"
    "model_name,seed,nodes,threads,processes,indices,actual_freq,P_hat,count,std_error,ci_lower,"
    "ci_upper";

int format_row(const mcr_csv_row* r, std::string* out) {
  char ts[32];
  const time_t t = (time_t)r->timestamp;
  struct tm tmv;
  if (localtime_r(&t, &tmv) == nullptr || strftime(ts, sizeof(ts), "%Y-%m-%d %H:%M:%S", &tmv) == 0)
    snprintf(ts, sizeof(ts), "UNKNOWN");
  std::string idx;
  for (int i = 0; i < r->num_indices; ++i) {
    if (!r->indices || !r->indices[i].name || !r->indices[i].file) {
      set_error("index %d has no name or file", i);
      return MCR_E_ARG;
    }
    if (i) idx.push_back(';');
    idx += r->indices[i].name;
    idx.push_back(':');
    idx += r->indices[i].file;
  }
  const std::string comment = csv_field(r->comment), indices = csv_field(idx.c_str());
  const int n = snprintf(nullptr, 0,
                         "%d,%s,%ld,%s,%s,%s,%.4f,%d,%d,%.4f,%s,%lu,%d,%d,%d,%s,%.6f,%.6f,%d,%.6f,"
                         "%.6f,%.6f",
                         r->iteration_id, ts, r->execution_time_ms, comment.c_str(),
                         r->start_date ? r->start_date : "", r->end_date ? r->end_date : "",
                         r->train_ratio, r->M, r->k, r->x, r->model_name ? r->model_name : "",
                         r->seed, r->nodes > 0 ? r->nodes : 0, r->threads > 0 ? r->threads : 0,
                         r->processes > 0 ? r->processes : 0, indices.c_str(), r->actual_freq,
                         r->P_hat, r->count, r->std_error, r->ci_lower, r->ci_upper);
  std::vector<char> buf((size_t)n + 1);
  snprintf(buf.data(), buf.size(),
           "%d,%s,%ld,%s,%s,%s,%.4f,%d,%d,%.4f,%s,%lu,%d,%d,%d,%s,%.6f,%.6f,%d,%.6f,%.6f,%.6f",
           r->iteration_id, ts, r->execution_time_ms, comment.c_str(),
           r->start_date ? r->start_date : "", r->end_date ? r->end_date : "", r->train_ratio,
           r->M, r->k, r->x, r->model_name ? r->model_name : "", r->seed,
           r->nodes > 0 ? r->nodes : 0, r->threads > 0 ? r->threads : 0,
           r->processes > 0 ? r->processes : 0, indices.c_str(), r->actual_freq, r->P_hat,
           r->count, r->std_error, r->ci_lower, r->ci_upper);
  *out = std::string(buf.data(), (size_t)n);
  return MCR_OK;
}

int copy_out(const std::string& s, char* buf, int cap) {
  if (!buf || cap <= (int)s.size()) {
    set_error("buffer of %d bytes cannot hold %zu", cap, s.size() + 1);
    return MCR_E_ARG;
  }
  memcpy(buf, s.c_str(), s.size() + 1);
  return (int)s.size();
}

}  // namespace

extern "C" {

void mcr_free(void* p) { free(p); }

int mcr_read_mu_bin(const char* path, double** mu_out, int* n_out) {
  if (!path || !mu_out || !n_out) {
    set_error("path, mu_out and n_out are required");
    return MCR_E_ARG;
  }
  File in;
  in.f = fopen(path, "rb");
  if (!in.f) {
    set_error("cannot open %s: %s", path, strerror(errno));
    return MCR_E_ARG;
  }
  int32_t n = 0;
  if (fread(&n, sizeof(n), 1, in.f) != 1) {
    set_error("cannot read the dimension from %s", path);
    return MCR_E_ARG;
  }
  if (n <= 0 || n > 1000000) {
    set_error("invalid dimension N=%d in %s (must be between 1 and 1000000)", n, path);
    return MCR_E_ARG;
  }
  double* mu = (double*)malloc((size_t)n * sizeof(double));
  if (!mu) {
    set_error("cannot allocate %d doubles", n);
    return MCR_E_NOMEM;
  }
  if (fread(mu, sizeof(double), (size_t)n, in.f) != (size_t)n) {
    free(mu);
    set_error("%s holds fewer than N=%d values", path, n);
    return MCR_E_ARG;
  }
  *mu_out = mu;
  *n_out = n;
  return MCR_OK;
}

int mcr_read_sigma_bin(const char* path, double** sigma_out, int* rows_out, int* cols_out) {
  if (!path || !sigma_out || !rows_out || !cols_out) {
    set_error("path and outputs are required");
    return MCR_E_ARG;
  }
  File in;
  in.f = fopen(path, "rb");
  if (!in.f) {
    set_error("cannot open %s: %s", path, strerror(errno));
    return MCR_E_ARG;
  }
  int32_t rc[2] = {0, 0};
  if (fread(rc, sizeof(int32_t), 2, in.f) != 2) {
    set_error("cannot read the dimensions from %s", path);
    return MCR_E_ARG;
  }
  if (rc[0] <= 0 || rc[1] <= 0) {
    set_error("invalid dimensions rows=%d, cols=%d in %s (must be positive)", rc[0], rc[1], path);
    return MCR_E_ARG;
  }
  const long long total = (long long)rc[0] * (long long)rc[1];
  if (total > 2000000000LL) {
    set_error("matrix too large (%lld elements) in %s", total, path);
    return MCR_E_ARG;
  }
  double* s = (double*)malloc((size_t)total * sizeof(double));
  if (!s) {
    set_error("cannot allocate %lld doubles", total);
    return MCR_E_NOMEM;
  }
  if (fread(s, sizeof(double), (size_t)total, in.f) != (size_t)total) {
    free(s);
    set_error("%s holds fewer than %lld values", path, total);
    return MCR_E_ARG;
  }
  *sigma_out = s;
  *rows_out = rc[0];
  *cols_out = rc[1];
  return MCR_OK;
}

int mcr_read_params_bin(const char* path, double* actual_freq_out) {
  if (!path || !actual_freq_out) {
    set_error("path and actual_freq_out are required");
    return MCR_E_ARG;
  }
  File in;
  in.f = fopen(path, "rb");
  if (!in.f) {
    set_error("cannot open %s: %s", path, strerror(errno));
    return MCR_E_ARG;
  }
  double f = 0.0;
  if (fread(&f, sizeof(double), 1, in.f) != 1) {
    set_error("cannot read actual_freq from %s", path);
    return MCR_E_ARG;
  }
  // the reference's range test: a NaN passes it
  if (f < 0.0 || f > 1.0) {
    set_error("invalid actual_freq=%.6f in %s (must be between 0.0 and 1.0)", f, path);
    return MCR_E_ARG;
  }
  *actual_freq_out = f;
  return MCR_OK;
}

int mcr_write_bins(const char* dir, const double* mu, int n, const double* sigma,
                   double actual_freq) {
  if (!dir || !mu || !sigma || n <= 0) {
    set_error("dir, mu, sigma and n >= 1 are required");
    return MCR_E_ARG;
  }
  MCR_TRY(make_dirs(dir));
  const std::string d(dir);
  const int32_t n32 = n;
  const int32_t rc[2] = {n, n};
  {
    const void* parts[2] = {&n32, mu};
    const size_t sizes[2] = {sizeof(n32), (size_t)n * sizeof(double)};
    MCR_TRY(write_file(d + "/mu.bin", parts, sizes, 2));
  }
  {
    const void* parts[2] = {rc, sigma};
    const size_t sizes[2] = {sizeof(rc), (size_t)n * n * sizeof(double)};
    MCR_TRY(write_file(d + "/sigma.bin", parts, sizes, 2));
  }
  {
    const void* parts[1] = {&actual_freq};
    const size_t sizes[1] = {sizeof(double)};
    MCR_TRY(write_file(d + "/params.bin", parts, sizes, 1));
  }
  return MCR_OK;
}

int mcr_load_config(const char* path, mcr_config* c) {
  if (!path || !c) {
    set_error("path and config are required");
    return MCR_E_ARG;
  }
  memset(c, 0, sizeof(*c));
  c->x = 0.02;
  c->N = 0;
  c->M = 100000;
  c->k = 5;
  c->random_seed = 42;
  c->train_ratio = 0.8;
  File in;
  in.f = fopen(path, "r");
  if (!in.f) {
    set_error("cannot open config file %s: %s", path, strerror(errno));
    return MCR_E_ARG;
  }
  std::vector<std::pair<int, std::string>> lines;  // (line number, text)
  {
    std::string cur;
    int ch, no = 1;
    while ((ch = fgetc(in.f)) != EOF) {
      if (ch == '\n') {
        lines.push_back({no++, cur});
        cur.clear();
      } else {
        cur.push_back((char)ch);
      }
    }
    if (!cur.empty()) lines.push_back({no, cur});
  }
  auto fail = [&](int no, const char* what) {
    set_error("%s:%d: %s", path, no, what);
    mcr_free_config(c);
    return MCR_E_ARG;
  };
  std::string block;  // key of the block being read ("" = none)
  bool block_has_items = false;
  for (auto& ln : lines) {
    const std::string raw = strip_comment(ln.second);
    if (trim(raw).empty()) continue;
    size_t ind = 0;
    while (ind < raw.size() && raw[ind] == ' ') ++ind;
    if (ind < raw.size() && raw[ind] == '\t') return fail(ln.first, "tab in indentation");
    const std::string body = trim(raw);
    if (ind == 0 && (body == "---" || body == "...")) continue;
    if (ind > 0) {
      if (block.empty()) return fail(ln.first, "indented line outside a block");
      block_has_items = true;
      if (body[0] == '-' && (body.size() == 1 || body[1] == ' ')) {
        std::string v;
        if (!unquote(trim(body.substr(1)), &v)) return fail(ln.first, "unterminated quote");
        if (block == "models") add_model(c, v);
      } else {
        const size_t col = key_colon(body);
        if (col == std::string::npos) return fail(ln.first, "expected 'key: value'");
        std::string k, v;
        if (!unquote(trim(body.substr(0, col)), &k) || !unquote(trim(body.substr(col + 1)), &v))
          return fail(ln.first, "unterminated quote");
        if (block == "indices") add_index(c, k, v);
      }
      continue;
    }
    // top level: close the previous block (an empty block is a null scalar)
    if (!block.empty() && !block_has_items) assign_scalar(c, block, "");
    block.clear();
    const size_t col = key_colon(body);
    if (col == std::string::npos) return fail(ln.first, "expected 'key: value'");
    std::string key;
    if (!unquote(trim(body.substr(0, col)), &key)) return fail(ln.first, "unterminated quote");
    const std::string val = trim(body.substr(col + 1));
    if (val.empty()) {
      block = key;
      block_has_items = false;
      clear_list(c, key);
      continue;
    }
    if (val.front() == '[' || val.front() == '{') {
      const char close = val.front() == '[' ? ']' : '}';
      if (val.back() != close) return fail(ln.first, "unterminated flow collection");
      clear_list(c, key);
      for (const std::string& item : split_flow(val.substr(1, val.size() - 2))) {
        if (close == ']') {
          std::string v;
          if (!unquote(item, &v)) return fail(ln.first, "unterminated quote");
          if (key == "models") add_model(c, v);
        } else {
          const size_t ic = key_colon(item);
          if (ic == std::string::npos) return fail(ln.first, "expected 'key: value' in mapping");
          std::string k, v;
          if (!unquote(trim(item.substr(0, ic)), &k) || !unquote(trim(item.substr(ic + 1)), &v))
            return fail(ln.first, "unterminated quote");
          if (key == "indices") add_index(c, k, v);
        }
      }
      continue;
    }
    std::string v;
    if (!unquote(val, &v)) return fail(ln.first, "unterminated quote");
    assign_scalar(c, key, v);
  }
  if (!block.empty() && !block_has_items) assign_scalar(c, block, "");
  return MCR_OK;
}

void mcr_free_config(mcr_config* c) {
  if (!c) return;
  free(c->start);
  free(c->end);
  free(c->comment);
  c->start = c->end = c->comment = nullptr;
  clear_list(c, "models");
  clear_list(c, "indices");
}

int mcr_csv_header(char* buf, int cap) { return copy_out(kCsvHeader, buf, cap); }

int mcr_csv_format_row(const mcr_csv_row* row, char* buf, int cap) {
  if (!row) {
    set_error("row is NULL");
    return MCR_E_ARG;
  }
  std::string s;
  MCR_TRY(format_row(row, &s));
  return copy_out(s, buf, cap);
}

int mcr_append_csv(const char* path, const mcr_csv_row* row) {
  if (!path || !row) {
    set_error("path and row are required");
    return MCR_E_ARG;
  }
  std::string line;
  MCR_TRY(format_row(row, &line));
  const std::string p(path);
  const size_t slash = p.rfind('/');
  if (slash != std::string::npos && slash > 0) MCR_TRY(make_dirs(p.substr(0, slash)));
  std::lock_guard<std::mutex> lk(g_csv_mu);
  const bool exists = access(path, F_OK) == 0;
  File out;
  out.f = fopen(path, "a");
  if (!out.f) {
    set_error("cannot open CSV file %s for writing: %s", path, strerror(errno));
    return MCR_E_ARG;
  }
  if (!exists && fprintf(out.f, "%s\n", kCsvHeader) < 0) {
    set_error("cannot write the CSV header to %s", path);
    return MCR_E_ARG;
  }
  if (fprintf(out.f, "%s\n", line.c_str()) < 0 || fflush(out.f) != 0) {
    set_error("cannot write to %s: %s", path, strerror(errno));
    return MCR_E_ARG;
  }
  const int rc = fclose(out.f);
  out.f = nullptr;
  if (rc != 0) {
    set_error("cannot close %s: %s", path, strerror(errno));
    return MCR_E_ARG;
  }
  return MCR_OK;
}

}  // extern "C"
