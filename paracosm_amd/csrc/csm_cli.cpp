// Drop-in command line for the MI355X engine: the flag set and report lines
// of the reference driver (ParaCOSM/CSM/matching/main.cpp:677-875) over the
// C-ABI in include/csm_c.h.
//
//   csm -q QUERY -d DATA -u STREAM [-a parallel_graphflow|parallel_symbi|...]
//       [--max-results N] [--time-limit S] [--print-prep 0|1] [--print-enum 0|1]
//       [--homo 0|1] [--report-initial 0|1] [--initial-time-limit S]
//       [-t N] [--auto-tuning 0|1]
//       [--compat 0|1] [--batch N] [--device N] [--reference-orders 0|1]
//       [--enum-cap N]   (rows kept per call with --print-enum; default 2^20)
//
// Default mode applies every update of the stream, in order, batched on the
// device (--batch, default 65534), insertions and deletions alike, and
// reports exact totals.  --compat 1 replays the reference's window loop
// (BatchUpdates3, main.cpp:253-397): windows of 16 updates, the first unsafe
// edge insertion of a window applied and searched, safe updates and edge
// deletions skipped, so that its totals can be compared line for line.
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "csm_c.h"

namespace {

constexpr uint32_t ABSENT = 0xFFFFFFFFu;

struct TextGraph {
  std::vector<uint32_t> vl, src, dst, el;
};

bool file_exists(const std::string& p) {
  std::ifstream f(p);
  return f.good();
}

bool read_file(const std::string& path, std::vector<char>& buf) {
  FILE* f = fopen(path.c_str(), "rb");
  if (!f) return false;
  buf.clear();
  char tmp[1 << 16];
  size_t got;
  while ((got = fread(tmp, 1, sizeof tmp, f)) > 0) buf.insert(buf.end(), tmp, tmp + got);
  fclose(f);
  return true;
}

// The whole file parsed on the device (csm_parse_text) in the grammar of
// Graph::LoadFromFile (graph.cpp:217-249) or Graph::LoadUpdateStream
// (graph.cpp:251-285).
bool parse_file(const std::string& path, int mode, int device, std::vector<csm_update>& recs) {
  std::vector<char> text;
  if (!file_exists(path) || !read_file(path, text)) {
    std::cout << "Failed to open: " << path << std::endl;
    return false;
  }
  size_t lines = 1;
  for (char c : text) lines += c == '\n';
  recs.resize(lines);
  size_t n = 0;
  uint64_t bad = 0;
  int rc = csm_parse_text(device, mode, text.data(), text.size(), recs.data(), recs.size(), &n, &bad);
  if (rc == CSM_E_ARG && bad) {
    std::cout << "malformed line " << bad << " in " << path << std::endl;
    return false;
  }
  if (rc) {
    std::cout << "parsing " << path << " on the device failed (" << rc << ")" << std::endl;
    return false;
  }
  recs.resize(n);
  return true;
}

// Graph records as Graph::AddVertex / AddEdge keep them (graph.cpp:23-93): a
// vertex keeps its first label; edges in file order.
bool load_graph(const std::string& path, int device, TextGraph& g) {
  std::vector<csm_update> recs;
  if (!parse_file(path, CSM_PARSE_GRAPH, device, recs)) return false;
  for (const csm_update& u : recs) {
    if (u.type == 'v') {
      if (u.id1 >= g.vl.size()) g.vl.resize(u.id1 + 1, ABSENT);
      if (g.vl[u.id1] == ABSENT) g.vl[u.id1] = u.label;
    } else {
      g.src.push_back(u.id1);
      g.dst.push_back(u.id2);
      g.el.push_back(u.label);
    }
  }
  return true;
}

bool load_stream(const std::string& path, int device, std::vector<csm_update>& ups) {
  return parse_file(path, CSM_PARSE_STREAM, device, ups);
}

size_t count_unique_edges(const TextGraph& g) {
  // as Graph::AddEdge ignores duplicates (graph.cpp:63-64)
  std::vector<uint64_t> k;
  k.reserve(g.src.size());
  for (size_t i = 0; i < g.src.size(); ++i) {
    uint32_t a = g.src[i], b = g.dst[i];
    if (a > b) std::swap(a, b);
    k.push_back(((uint64_t)a << 32) | b);
  }
  std::sort(k.begin(), k.end());
  return std::unique(k.begin(), k.end()) - k.begin();
}

long peak_vm_kb() {  // utils.h:38-67
  FILE* f = fopen("/proc/self/status", "r");
  if (!f) return -1;
  char line[256];
  long v = -1;
  while (fgets(line, sizeof line, f))
    if (strncmp(line, "VmPeak:", 7) == 0) {
      v = atol(line + 7);
      break;
    }
  fclose(f);
  return v;
}

double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

csm_graph as_graph(const TextGraph& g) {
  csm_graph c;
  c.num_vertices = (uint32_t)g.vl.size();
  c.vlabels = g.vl.data();
  c.num_edges = g.src.size();
  c.src = g.src.data();
  c.dst = g.dst.data();
  c.elabels = g.el.data();
  return c;
}

int die(csm_engine* e, int rc) {
  if (rc == CSM_E_DELETE_MISSING) std::cout << "deletion error" << std::endl;
  else if (e) std::cout << "error: " << csm_last_error(e) << std::endl;
  else std::cout << "error " << rc << std::endl;
  if (e) csm_destroy(e);
  return -1;
}

}  // namespace

// --print-enum: the embeddings of the last call, one line each with the data
// vertex of every query vertex followed by a space (the reference prints `m`
// so as each search reaches it, SymBi/symbi.cpp:564-569), grouped by update
// in stream order and sorted within an update, so that the output does not
// depend on the device's search order.
static void print_embeddings(csm_engine* e) {
  size_t n = 0;
  uint64_t found = 0;
  uint32_t w = 0;
  if (csm_embeddings(e, nullptr, 0, &n, &found, &w) || w == 0) return;
  std::vector<uint32_t> rows((size_t)n * w);
  if (n && csm_embeddings(e, rows.data(), n, &n, &found, &w)) return;
  std::vector<size_t> idx(n);
  for (size_t r = 0; r < n; ++r) idx[r] = r;
  std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    return std::lexicographical_compare(&rows[a * w], &rows[a * w] + w, &rows[b * w],
                                        &rows[b * w] + w);
  });
  for (size_t r : idx) {
    const uint32_t* x = &rows[r * w];
    for (uint32_t i = 2; i < w && x[i] != 0xFFFFFFFFu; ++i) std::cout << x[i] << " ";
    std::cout << "\n";
  }
  if (found > n) std::cout << found - n << " more embeddings not kept (--enum-cap)\n";
}

int main(int argc, char** argv) {
  std::string query_path, data_path, stream_path, algorithm = "none";
  uint64_t max_results = 0;
  double time_limit = 0.0, initial_time_limit = 0.0;
  int print_prep = 1, print_enum = 0, homo = 0, report_initial = 1, compat = 0;
  uint32_t threads = 6, auto_tuning = 1, batch = 0, device = 0, reference_orders = 0;
  uint32_t enum_cap = 0;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() -> std::string {
      if (i + 1 >= argc) { std::cerr << a << " needs a value\n"; exit(1); }
      return argv[++i];
    };
    if (a == "-q" || a == "--query") query_path = next();
    else if (a == "-d" || a == "--data") data_path = next();
    else if (a == "-u" || a == "--update") stream_path = next();
    else if (a == "-a" || a == "--algorithm") algorithm = next();
    else if (a == "--max-results") max_results = std::stoull(next());
    else if (a == "--time-limit") time_limit = std::stod(next());
    else if (a == "--print-prep") print_prep = std::stoi(next());
    else if (a == "--print-enum") print_enum = std::stoi(next());
    else if (a == "--enum-cap") enum_cap = (uint32_t)std::stoul(next());
    else if (a == "--homo") homo = std::stoi(next());
    else if (a == "--report-initial") report_initial = std::stoi(next());
    else if (a == "--initial-time-limit") initial_time_limit = std::stod(next());
    else if (a == "--orders") next();  // accepted; orders are planned on the host
    else if (a == "--reference-orders") reference_orders = (uint32_t)std::stoul(next());
    else if (a == "-t" || a == "--thread-num") threads = (uint32_t)std::stoul(next());
    else if (a == "--auto-tuning") auto_tuning = (uint32_t)std::stoul(next());
    else if (a == "--compat") compat = std::stoi(next());
    else if (a == "--batch") batch = (uint32_t)std::stoul(next());
    else if (a == "--device") device = (uint32_t)std::stoul(next());
    else { std::cerr << "unknown option " << a << "\n"; return 1; }
  }
  if (query_path.empty() || data_path.empty() || stream_path.empty()) {
    std::cerr << "-q, -d and -u are required\n";
    return 1;
  }

  auto t0 = std::chrono::steady_clock::now();
  std::cout << "----------- Loading graphs ------------" << std::endl;
  TextGraph q;
  if (!load_graph(query_path, (int)device, q)) return -1;
  std::cout << "# vertices = " << q.vl.size() << "\n# edges = " << count_unique_edges(q) << std::endl;
  csm_opts o;
  memset(&o, 0, sizeof o);
  o.max_results = max_results == UINT32_MAX ? 0 : max_results;
  o.print_prep = (uint32_t)print_prep;
  o.print_enum = (uint32_t)print_enum;
  o.enum_cap = enum_cap;
  o.homo = (uint32_t)homo;
  o.threads = threads;
  o.auto_tuning = auto_tuning;
  o.max_batch = batch;
  o.time_limit_s = 0;   // set per phase below
  o.device = device;
  o.reference_orders = reference_orders;
  // the data graph: its text parsed on the device into records that stay
  // there, and the engine's CSR built from them on the device (the
  // reference's LoadFromFile, graph.cpp:217-249)
  csm_graph cq = as_graph(q);
  std::vector<char> text;
  if (!file_exists(data_path) || !read_file(data_path, text)) {
    std::cout << "Failed to open: " << data_path << std::endl;
    return -1;
  }
  csm_update* d_rec = nullptr;
  size_t n_rec = 0;
  uint64_t bad = 0;
  int rc = csm_parse_text_to_device((int)device, CSM_PARSE_GRAPH, text.data(), text.size(), &d_rec,
                                    &n_rec, &bad);
  if (rc == CSM_E_ARG && bad) {
    std::cout << "malformed line " << bad << " in " << data_path << std::endl;
    return -1;
  }
  if (rc) {
    std::cout << "parsing " << data_path << " on the device failed (" << rc << ")" << std::endl;
    return -1;
  }
  int err = 0;
  csm_engine* e = csm_create_from_records(algorithm.c_str(), &cq, d_rec, n_rec, &o, &err);
  csm_free_records((int)device, d_rec);
  if (!e) {
    if (err == CSM_E_UNKNOWN_ALGO) std::cout << "Unknown algorithm" << std::endl;
    else std::cout << "engine creation failed (" << err << ")" << std::endl;
    return -1;
  }
  {
    size_t de = 0, dv = 0;
    csm_memory(e, &de, &dv);
    std::cout << "# vertices = " << dv << "\n# edges = " << de << std::endl;
  }
  std::cout << "Load Graphs: " << ms_since(t0) << "ms" << std::endl;

  std::cout << "------------ Preprocessing ------------" << std::endl;
  t0 = std::chrono::steady_clock::now();
  rc = csm_preprocess(e);
  if (rc) return die(e, rc);
  std::cout << "Preprocessing: " << ms_since(t0) << "ms" << std::endl;

  if (report_initial) {
    std::cout << "----------- Initial Matching ----------" << std::endl;
    t0 = std::chrono::steady_clock::now();
    uint64_t n0 = 0;
    csm_set_time_limit(e, initial_time_limit);
    rc = csm_initial(e, &n0);
    const bool timeout = rc == CSM_E_TIME_LIMIT;
    if (timeout) std::cout << "Timeout " << initial_time_limit << "s" << std::endl;
    else if (rc) return die(e, rc);
    if (print_enum) print_embeddings(e);
    std::cout << "Initial Matching: " << ms_since(t0) << "ms" << std::endl;
    std::cout << n0 << " matches. \n";
    // main.cpp:779-790: a timed-out initial matching ends the run
    if (timeout) {
      csm_destroy(e);
      return 1;
    }
  }

  std::cout << "--------- Incremental Matching --------" << std::endl;
  if (auto_tuning == 1) std::cout << "auto_tuning on:" << auto_tuning << std::endl;
  else {
    std::cout << "auto_tuning off:" << auto_tuning << std::endl;
    std::cout << "use " << threads << " threads for incresemental matching " << std::endl;
  }
  std::vector<csm_update> ups;
  if (!load_stream(stream_path, (int)device, ups)) {
    csm_destroy(e);
    return -1;
  }
  size_t num_v = 0, num_e = 0, unsafe = 0, classified_unsafe = 0, partial = 0;
  t0 = std::chrono::steady_clock::now();
  if (compat) {
    // BatchUpdates3 (main.cpp:253-397)
    const size_t W = 16, n = ups.size();
    size_t base = 0;
    uint64_t last_pos = 0, last_neg = 0;
    while (base < n) {
      bool safe_all = true;
      size_t min_safe = W + 10;
      for (size_t i = 0; i < W; ++i) {
        if (i + base >= n) continue;
        const csm_update& u = ups[i + base];
        int s = 1;
        if (u.type == 'e') {
          if ((rc = csm_classify(e, u.id1, u.id2, u.label, &s))) return die(e, rc);
        } else if (u.is_add) {
          if ((rc = csm_add_vertex(e, u.id1, u.label))) return die(e, rc);
          continue;
        } else {
          s = 0;
        }
        if (!s) {
          safe_all = false;
          if (min_safe > i) min_safe = i;
          break;
        }
      }
      if (!safe_all) {
        const csm_update& u = ups[min_safe + base];
        ++classified_unsafe;
        if (u.type == 'e' && u.is_add) {
          uint64_t p = 0;
          if ((rc = csm_add_edge(e, u.id1, u.id2, u.label, &p))) return die(e, rc);
          if (print_enum) print_embeddings(e);
        } else if (u.type == 'v' && !u.is_add) {
          if ((rc = csm_remove_vertex(e, u.id1))) return die(e, rc);
        }
        base += min_safe + 1;
        num_e += min_safe + 1;
        uint64_t pc = 0, nc = 0;
        csm_totals(e, nullptr, &pc, &nc);
        if (pc != last_pos || nc != last_neg) {
          last_pos = pc;
          last_neg = nc;
          ++unsafe;
        }
      } else {
        base += W;
        num_e += W;
      }
    }
  } else {
    const size_t B = batch ? batch : 65534;
    std::vector<uint64_t> pos, neg;
    std::vector<uint8_t> safe, status;
    const size_t nq = csm_num_queries(e);
    // --time-limit covers the whole incremental phase: each call gets what
    // is left; once it is spent every later search stops at once while the
    // updates are still applied (the reference's reach_time_limit polls)
    const auto t_end = t0 + std::chrono::duration<double>(time_limit);
    bool timeout = false;
    for (size_t s = 0; s < ups.size(); s += B) {
      size_t m = std::min(B, ups.size() - s);
      pos.assign(m * nq, 0);
      neg.assign(m * nq, 0);
      safe.assign(m, 0);
      status.assign(m, 0);
      if (time_limit > 0) {
        double left = std::chrono::duration<double>(t_end - std::chrono::steady_clock::now()).count();
        csm_set_time_limit(e, left > 1e-9 ? left : 1e-9);
      }
      rc = csm_apply_batch(e, ups.data() + s, m, pos.data(), neg.data(), safe.data(), nullptr);
      if (rc == CSM_E_TIME_LIMIT) {
        if (!timeout) std::cout << "Timeout " << time_limit << "s" << std::endl;
        timeout = true;
      } else if (rc) {
        return die(e, rc);
      }
      if (print_enum) print_embeddings(e);
      csm_update_status(e, status.data(), m);
      for (size_t i = 0; i < m; ++i) partial += (status[i] & CSM_USTAT_TIME_LIMIT) != 0;
      for (size_t i = 0; i < m; ++i) {
        const csm_update& u = ups[s + i];
        if (u.type == 'v') ++num_v; else ++num_e;
        uint64_t c = 0;
        for (size_t j = 0; j < nq; ++j) c += pos[i * nq + j] + neg[i * nq + j];
        unsafe += c > 0;
        classified_unsafe += !safe[i];
      }
    }
  }
  std::cout << "Incremental Matching: " << ms_since(t0) << "ms" << std::endl;
  std::cout << num_v << " vertex updates.\n";
  std::cout << num_e << " edge updates.\n";
  std::cout << unsafe << " unsafe updates.\n";
  uint64_t ptot = 0, ntot = 0;
  csm_totals(e, nullptr, &ptot, &ntot);
  std::cout << ptot << " positive matches.\n";
  std::cout << ntot << " negative matches.\n";
  std::cout << classified_unsafe << " updates classified unsafe by the label filter.\n";
  if (time_limit > 0 && !compat)
    std::cout << partial << " updates with partial counts (time limit).\n";
  size_t me = 0, mv = 0;
  csm_memory(e, &me, &mv);
  uint64_t dev = 0;
  csm_device_bytes(e, &dev);
  std::cout << "\n# edges in index in total: " << me;
  std::cout << "\n# vertices in index in total: " << mv;
  std::cout << "\nDevice memory: " << dev / 1024 << " KB";
  std::cout << "\nPeak Virtual Memory: " << peak_vm_kb() << " KB";
  std::cout << "\n\n----------------- End -----------------" << std::endl;
  csm_destroy(e);
  return 0;
}
