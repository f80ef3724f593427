// knnjoin -- command-line driver with the reference's interface.
//
//   knnjoin query_nb source_nb dim qrep_nb srep_nb K qfile sfile [options]
//
// Positional arguments as the reference's main() (src/knnjoin.cu:990-1001):
// counts, dimension, representative counts per side, k, and two input
// files, each a path or the keyword "random".  Differences, all deliberate:
//   - the text loader is row-aware: line i is row i, its first `dim`
//     whitespace/comma separated numbers are the coordinates and further
//     columns are ignored (the reference reads fscanf("%f") tokens across
//     lines, include/common8.h:109-112, 130-132, so a labelled file only
//     parses at the column count); files ending in .f32 are raw row-major
//     binary32;
//   - "random" reproduces the reference generator (srand(2015) once, then
//     rand() % 10 + (float)rand() / RAND_MAX for every query coordinate,
//     then every source coordinate; common8.h:94-123);
//   - the same query/source file and counts make a self-join;
//   - ties break by (squared distance, row), and the result is exact.
// Output: stage times, then the reference's lines "i,k %d %d  %d %f" for
// query 100 (knnjoin.cu:1143-1145; the last query if there are fewer).
//
// Options:
//   --precision auto|fp32|bf16|bf16x1   distance filter (default auto)
//   --kernel auto|valu|mfma
//   --query I          print query I instead of 100 (-1: none)
//   --out FILE         write every result as binary IndexDist[query_nb][K]
//                      ({int32 index; float32 dist}, include/common8.h:37-40)
//   --device N
//   --dump-input N     print the first N parsed rows of each set and exit
//                      (no GPU needed)
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/knnj.h"

namespace {

[[noreturn]] void die(const std::string &msg, int code = 1) {
  std::fprintf(stderr, "knnjoin: %s\n", msg.c_str());
  std::exit(code);
}

bool ends_with(const std::string &s, const char *suf) {
  const size_t n = std::strlen(suf);
  return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

// Row-aware text or raw binary32 rows into out[n][d].
void load_rows(const std::string &path, long long n, int d, float *out) {
  if (ends_with(path, ".f32")) {
    std::ifstream f(path, std::ios::binary);
    if (!f) die("cannot open " + path);
    f.read(reinterpret_cast<char *>(out), (std::streamsize)(sizeof(float) * n * d));
    if (f.gcount() != (std::streamsize)(sizeof(float) * n * d))
      die(path + ": fewer than " + std::to_string(n) + " rows of " + std::to_string(d) + " floats");
    return;
  }
  std::ifstream f(path);
  if (!f) die("cannot open " + path);
  std::string line;
  long long row = 0, lineno = 0;
  while (row < n && std::getline(f, line)) {
    ++lineno;
    for (char &c : line)
      if (c == ',' || c == '\t' || c == '\r' || c == ';') c = ' ';
    const char *p = line.c_str();
    char *end = nullptr;
    int got = 0;
    while (got < d) {
      while (*p == ' ') ++p;
      if (!*p) break;
      const float v = std::strtof(p, &end);
      if (end == p) die(path + ":" + std::to_string(lineno) + ": not a number: " + std::string(p).substr(0, 20));
      out[row * d + got++] = v;
      p = end;
    }
    if (got == 0) continue;  // blank line
    if (got < d)
      die(path + ":" + std::to_string(lineno) + ": " + std::to_string(got) + " values, need " + std::to_string(d));
    ++row;
  }
  if (row < n) die(path + ": " + std::to_string(row) + " rows, need " + std::to_string(n));
}

const char *prec_name(int p) {
  switch (p) {
    case KNNJ_PREC_FP32: return "fp32";
    case KNNJ_PREC_BF16: return "bf16 hi/lo filter + binary64 rerank";
    case KNNJ_PREC_BF16X1: return "bf16 filter + binary64 rerank";
  }
  return "?";
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 9) {
    std::fprintf(stderr,
                 "usage: %s query_nb source_nb dimension qrep_nb srep_nb k query_file source_file [options]\n"
                 "  files: text (one row per line) | raw binary32 (*.f32) | random\n"
                 "  options: --precision auto|fp32|bf16|bf16x1  --kernel auto|valu|mfma  --query I\n"
                 "           --out FILE  --device N  --dump-input N\n",
                 argv[0]);
    return 1;
  }
  const long long query_nb = std::atoll(argv[1]), source_nb = std::atoll(argv[2]);
  const int dim = std::atoi(argv[3]), qrep_nb = std::atoi(argv[4]), srep_nb = std::atoi(argv[5]);
  const int K = std::atoi(argv[6]);
  const std::string qfile = argv[7], sfile = argv[8];
  if (query_nb < 1 || source_nb < 1 || dim < 1 || K < 1 || qrep_nb < 0 || srep_nb < 0)
    die("counts, dimension and k must be positive");
  knnj_opts o{};
  o.m_q = qrep_nb;
  o.m_r = srep_nb;
  long long show = 100;
  std::string out_path;
  int device = 0;
  long long dump = -1;
  for (int i = 9; i < argc; ++i) {
    const std::string a = argv[i];
    auto val = [&]() -> std::string {
      if (i + 1 >= argc) die("missing value after " + a);
      return argv[++i];
    };
    if (a == "--precision") {
      const std::string v = val();
      o.precision = v == "fp32" ? KNNJ_PREC_FP32 : v == "bf16" ? KNNJ_PREC_BF16 : v == "bf16x1" ? KNNJ_PREC_BF16X1
                    : v == "auto" ? KNNJ_PREC_AUTO : -1;
      if (o.precision < 0) die("unknown precision " + v);
    } else if (a == "--kernel") {
      const std::string v = val();
      o.kernel = v == "valu" ? KNNJ_KERNEL_VALU : v == "mfma" ? KNNJ_KERNEL_MFMA : v == "auto" ? KNNJ_KERNEL_AUTO : -1;
      if (o.kernel < 0) die("unknown kernel " + v);
    } else if (a == "--query") {
      show = std::atoll(val().c_str());
    } else if (a == "--out") {
      out_path = val();
    } else if (a == "--device") {
      device = std::atoi(val().c_str());
    } else if (a == "--dump-input") {
      dump = std::atoll(val().c_str());
    } else {
      die("unknown option " + a);
    }
  }

  const bool self = qfile == sfile && query_nb == source_nb && qfile != "random";
  std::vector<float> queries((size_t)query_nb * dim), sources;
  if (!self) sources.resize((size_t)source_nb * dim);
  // the reference seeds once and draws the queries first (common8.h:94-123)
  std::srand(2015);
  auto fill = [&](const std::string &f, long long n, float *x, const char *what) {
    if (f == "random") {
      std::printf("%s data random\n", what);
      for (long long i = 0; i < n * dim; ++i) {
        const int a = std::rand() % 10;
        x[i] = (float)a + (float)std::rand() / RAND_MAX;
      }
    } else {
      load_rows(f, n, dim, x);
    }
  };
  fill(qfile, query_nb, queries.data(), "query");
  if (!self) fill(sfile, source_nb, sources.data(), "source");
  const float *src = self ? queries.data() : sources.data();
  if (dump >= 0) {
    for (int side = 0; side < 2; ++side) {
      const float *x = side == 0 ? queries.data() : src;
      const long long n = std::min<long long>(dump, side == 0 ? query_nb : source_nb);
      for (long long i = 0; i < n; ++i) {
        std::printf("%s %lld:", side == 0 ? "query" : "source", i);
        for (int j = 0; j < dim; ++j) std::printf(" %.9g", x[i * dim + j]);
        std::printf("\n");
      }
    }
    return 0;
  }

  knnj_ctx *ctx = nullptr;
  if (knnj_ctx_create(device, &ctx) != KNNJ_OK) die(std::string("no device: ") + knnj_last_error(), 2);
  knnj_matrix Q{queries.data(), query_nb, dim, dim, 0};
  knnj_matrix R{src, source_nb, dim, dim, 0};
  o.out_aos = 1;
  std::vector<int32_t> knearest((size_t)query_nb * K * 2);
  knnj_stats st{};
  const auto t0 = std::chrono::steady_clock::now();
  const knnj_status e = knnj_join(ctx, R, Q, K, &o, knearest.data(), nullptr, &st);
  const auto t1 = std::chrono::steady_clock::now();
  if (e != KNNJ_OK) die(std::string("join failed: ") + knnj_last_error(), 3);
  std::printf("index time %f\n", st.t_index_ms * 1e-3);
  std::printf("bounds time %f\n", st.t_bounds_ms * 1e-3);
  std::printf("join time %f\n", (st.t_join_ms + st.t_fallback_ms) * 1e-3);
  std::printf("total time %f\n", std::chrono::duration<double>(t1 - t0).count());
  std::printf("precision %s, reps %d/%d, pruning ratio %.6f, distance evaluations %lld\n",
              prec_name(st.precision_used), st.m_q, st.m_r, st.pruning_ratio, (long long)st.n_dist_evals);
  if (show >= 0) {
    const long long i = std::min(show, query_nb - 1);
    for (int j = 0; j < K; ++j) {
      float dist;
      std::memcpy(&dist, &knearest[((size_t)i * K + j) * 2 + 1], sizeof(float));
      std::printf("i,k %lld %d  %d %f\n", i, j, knearest[((size_t)i * K + j) * 2], dist);
    }
  }
  if (!out_path.empty()) {
    std::ofstream f(out_path, std::ios::binary);
    if (!f) die("cannot write " + out_path);
    f.write(reinterpret_cast<const char *>(knearest.data()), (std::streamsize)(knearest.size() * sizeof(int32_t)));
  }
  knnj_ctx_destroy(ctx);
  return 0;
}
