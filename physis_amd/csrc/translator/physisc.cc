// physisc-hip: translates a Physis C program for the HIP target.
//
//   physisc-hip [--config FILE] [-I DIR] [-D NAME[=V]] prog.c [-o prog.hip.cpp]
//
// Mirrors the reference driver's flow (translator/physisc.cc:448-692):
// preprocess with PHYSIS_USER defined, build the program tree, translate,
// write <prog>.hip.cpp.  #include lines of the input are kept as written
// (the Physis header is replaced by the HIP target's); user macros are
// expanded by the C preprocessor.
//
// Configuration file: the reference's format (KEY = value per line, `--`
// comments, Lua tables whose first entry is used, as the reference does when
// not auto-tuning: translator/configuration.cc:27-60).  Keys:
//   HIP_BLOCK_WAVES, HIP_PREFETCH     launch shape of the run kernels
//   HIP_KERNEL_ERROR_CHECK, CUDA_KERNEL_ERROR_CHECK
//                                     wait and check after each run
//   TRACE_KERNEL                      trace output of run functions
//   HIP_TWO_SWEEP                     two-sweep fusion of map pairs (true)
//   OPT_LOOP_PEELING                  interior form of each wave step (true)
// and, accepted with no effect because the HIP run kernel always does what
// they ask (translator/optimizer/*.cc): OPT_KERNEL_INLINING,
// OPT_REGISTER_BLOCKING, OPT_UNCONDITIONAL_GET, OPT_OFFSET_CSE,
// OPT_OFFSET_SPATIAL_CSE, OPT_OFFSET_COMP, OPT_LOOP_OPT; CUDA_BLOCK_SIZE is
// a CUDA launch shape and is ignored.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include <unistd.h>

#include "parser.h"
#include "translator.h"

using namespace physisc;

namespace {

[[noreturn]] void Usage() {
  std::fprintf(stderr,
               "usage: physisc-hip [--config FILE] [-I DIR] [-D NAME[=V]] "
               "prog.c [-o out.hip.cpp]\n");
  std::exit(2);
}

std::string ReadFile(const std::string &p) {
  std::ifstream f(p);
  if (!f) {
    std::fprintf(stderr, "physisc-hip: cannot read %s\n", p.c_str());
    std::exit(1);
  }
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

std::string Trim(const std::string &s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  size_t b = s.find_last_not_of(" \t\r\n");
  return a == std::string::npos ? "" : s.substr(a, b - a + 1);
}

void LoadConfig(const std::string &path, Config *cfg) {
  std::istringstream is(ReadFile(path));
  std::string line;
  while (std::getline(is, line)) {
    size_t h = line.find_first_of("#-");
    if (h != std::string::npos && line.compare(h, 2, "--") == 0)
      line = line.substr(0, h);
    size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    std::string k = Trim(line.substr(0, eq)), v = Trim(line.substr(eq + 1));
    // a Lua table of alternatives: its first entry
    while (!v.empty() && v[0] == '{') {
      v = Trim(v.substr(1));
      size_t e = v.find_first_of(",}");
      if (v.empty() || v[0] == '{') continue;
      v = Trim(v.substr(0, e));
    }
    bool on = v == "true" || v == "1";
    if (k == "HIP_BLOCK_WAVES") cfg->block_waves = std::atoi(v.c_str());
    else if (k == "HIP_PREFETCH") cfg->prefetch = std::atoi(v.c_str());
    else if (k == "HIP_KERNEL_ERROR_CHECK" || k == "CUDA_KERNEL_ERROR_CHECK")
      cfg->kernel_error_check = on;
    else if (k == "TRACE_KERNEL") cfg->trace = on;
    else if (k == "HIP_TWO_SWEEP") cfg->two_sweep = on;
    else if (k == "OPT_LOOP_PEELING") cfg->interior_split = on;
    else if (k.compare(0, 4, "OPT_") == 0 || k == "CUDA_BLOCK_SIZE")
      continue;  // see the header comment
    else
      std::fprintf(stderr, "physisc-hip: %s: unknown configuration key %s\n",
                   path.c_str(), k.c_str());
  }
  if (cfg->prefetch != 0 && (cfg->prefetch < 2 || cfg->prefetch % 2))
    cfg->prefetch = 2;
  int bw = cfg->block_waves;
  if (bw != 0 && bw != 4 && bw != 8 && bw != 16) cfg->block_waves = 0;
}

}  // namespace

int main(int argc, char **argv) {
  std::string in, out, config;
  std::vector<std::string> cpp_args;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "-o" && i + 1 < argc) out = argv[++i];
    else if (a == "--config" && i + 1 < argc) config = argv[++i];
    else if ((a == "-I" || a == "-D") && i + 1 < argc) cpp_args.push_back(a + argv[++i]);
    else if (a.compare(0, 2, "-I") == 0 || a.compare(0, 2, "-D") == 0) cpp_args.push_back(a);
    else if (!a.empty() && a[0] == '-') Usage();
    else in = a;
  }
  if (in.empty()) Usage();
  if (out.empty()) {
    size_t dot = in.rfind('.');
    out = (dot == std::string::npos ? in : in.substr(0, dot)) + ".hip.cpp";
  }
  Config cfg;
  if (!config.empty()) LoadConfig(config, &cfg);

  // 1. Replace #include lines by markers the parser keeps in place.
  std::istringstream src(ReadFile(in));
  std::string line, body;
  std::map<std::string, std::string> includes;
  int nmark = 0;
  while (std::getline(src, line)) {
    std::string t = Trim(line);
    if (t.compare(0, 1, "#") == 0) {
      std::string d = Trim(t.substr(1));
      if (d.compare(0, 7, "include") == 0) {
        std::string marker = "__physisc_include_" + std::to_string(nmark++) + "__";
        if (d.find("physis/") == std::string::npos) includes[marker] = t;
        else includes[marker] = "";
        body += marker + "\n";
        continue;
      }
    }
    body += line + "\n";
  }

  // 2. Preprocess (user macros, conditionals) with PHYSIS_USER defined.
  char tmpl[] = "/tmp/physisc_hip_XXXXXX.c";
  int fd = mkstemps(tmpl, 2);
  if (fd < 0) {
    std::perror("physisc-hip: mkstemps");
    return 1;
  }
  if (write(fd, body.data(), body.size()) != (ssize_t)body.size()) {
    std::perror("physisc-hip: write");
    return 1;
  }
  close(fd);
  std::string cmd = "cpp -P -DPHYSIS_USER -DPHYSIS_HIP_TRANSLATION";
  for (auto &a : cpp_args) cmd += " '" + a + "'";
  cmd += std::string(" '") + tmpl + "'";
  FILE *p = popen(cmd.c_str(), "r");
  if (!p) {
    std::perror("physisc-hip: cpp");
    return 1;
  }
  std::string pre;
  char buf[65536];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, p)) > 0) pre.append(buf, n);
  int rc = pclose(p);
  unlink(tmpl);
  if (rc != 0) {
    std::fprintf(stderr, "physisc-hip: preprocessing %s failed\n", in.c_str());
    return 1;
  }

  // 3. Parse and translate.
  Parser parser(Lex(pre, in), in);
  Program prog = parser.ParseProgram();
  {
    size_t s = in.find_last_of('/');
    prog.file = s == std::string::npos ? in : in.substr(s + 1);
  }
  Translator tr(prog, cfg, parser.globals());
  std::string code = tr.Translate();

  // 4. Markers back to the original include lines.
  for (auto &m : includes) {
    size_t pos;
    while ((pos = code.find(m.first)) != std::string::npos)
      code.replace(pos, m.first.size(), m.second);
  }
  std::ofstream o(out);
  o << code;
  if (!o) {
    std::fprintf(stderr, "physisc-hip: cannot write %s\n", out.c_str());
    return 1;
  }
  return 0;
}
