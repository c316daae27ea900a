// physisc-hip: stencil analysis, host rewriting (HIPTranslator) and code
// emission (HIPRuntimeBuilder).  See translator.h.
#include "translator.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <sstream>

#include "parser.h"

namespace physisc {

namespace {

[[noreturn]] void Fail(const std::string &where, const std::string &msg) {
  std::fprintf(stderr, "physisc-hip: %s: %s\n", where.c_str(), msg.c_str());
  std::exit(1);
}

std::string Strip(const std::string &t) {
  // drop qualifiers and spaces: "const int" -> "int"
  std::istringstream is(t);
  std::string w, out;
  while (is >> w) {
    if (w == "const" || w == "volatile" || w == "register") continue;
    out += (out.empty() ? "" : " ") + w;
  }
  return out;
}

bool IsIntType(const std::string &t) {
  std::string s = Strip(t);
  return s == "int" || s == "long" || s == "PSIndex" || s == "unsigned" ||
         s == "unsigned int" || s == "long int" || s == "short" ||
         s == "int32_t" || s == "int64_t";
}

std::string PsTypeOf(const std::string &elem) {
  if (elem == "float") return "PS_FLOAT";
  if (elem == "double") return "PS_DOUBLE";
  if (elem == "int") return "PS_INT";
  return "PS_LONG";
}

std::string ReduceFnOf(const std::string &elem) {
  if (elem == "float") return "__PSReduceGridFloat";
  if (elem == "double") return "__PSReduceGridDouble";
  if (elem == "int") return "__PSReduceGridInt";
  return "__PSReduceGridLong";
}

std::string Cap(const std::string &elem) {
  if (elem == "float") return "Float";
  if (elem == "double") return "Double";
  if (elem == "int") return "Int";
  return "Long";
}

const std::string *CalleeName(const Expr &e) {
  if (e.kind == EK::kCall && e.kids[0]->kind == EK::kIdent)
    return &e.kids[0]->text;
  return nullptr;
}

ExprP Raw(const std::string &s) {
  auto e = std::make_unique<Expr>(EK::kRaw);
  e->text = s;
  return e;
}

// Integer constant expression: literal, (c), -c, +c.
bool ConstInt(const Expr &e, int *v) {
  if (e.kind == EK::kLiteral) {
    const std::string &t = e.text;
    if (t.empty() || !std::isdigit((unsigned char)t[0])) return false;
    if (t.find_first_of(".eEpP") != std::string::npos &&
        !(t.size() > 1 && (t[1] == 'x' || t[1] == 'X')))
      return false;
    *v = (int)std::strtol(t.c_str(), nullptr, 0);
    return true;
  }
  if (e.kind == EK::kParen) return ConstInt(*e.kids[0], v);
  if (e.kind == EK::kUnary && (e.op == "-" || e.op == "+")) {
    int w;
    if (!ConstInt(*e.kids[0], &w)) return false;
    *v = e.op == "-" ? -w : w;
    return true;
  }
  return false;
}

std::string Join(const std::vector<std::string> &v, const std::string &sep) {
  std::string s;
  for (size_t i = 0; i < v.size(); ++i) s += (i ? sep : "") + v[i];
  return s;
}

}  // namespace

Translator::Translator(
    Program &prog, const Config &cfg,
    const std::vector<std::pair<std::string, std::string>> &globals)
    : prog_(prog), cfg_(cfg) {
  for (const auto &g : globals) {
    std::istringstream is(g.first);
    std::string w;
    while (is >> w)
      if (IsGridTypeName(w, nullptr, nullptr)) grid_vars_[g.second] = w;
  }
}

// ================================================================ analysis
void Translator::FindKernelsIn(const Expr &e) {
  if (const std::string *c = CalleeName(e)) {
    int karg = -1;
    bool reduce = false;
    if (*c == "PSStencilMap" && e.kids.size() >= 3) karg = 1;
    if (*c == "PSStencilMapReduce" && e.kids.size() >= 5) {
      karg = 3;
      reduce = true;
    }
    if (*c == "PSReduce" && e.kids.size() >= 5) {  // kernel form
      karg = 3;
      reduce = true;
    }
    if (*c == "PSStencilMapRedBlack" || *c == "PSStencilMapRed" ||
        *c == "PSStencilMapBlack")
      Fail(prog_.file, "red-black stencil maps are not supported by the HIP "
                       "target yet");
    if (karg > 0) {
      const Expr *k = e.kids[karg].get();
      while (k->kind == EK::kParen || k->kind == EK::kCast) k = k->kids[0].get();
      if (k->kind != EK::kIdent)
        Fail(prog_.file, *c + ": the kernel must be a function name");
      const std::string &name = k->text;
      if (!kernels_.count(name)) {
        KernelInfo ki;
        ki.name = name;
        for (auto &it : prog_.items)
          if (it.kind == TopItem::kFunction && it.func->name == name &&
              it.func->body)
            ki.fn = it.func.get();
        if (!ki.fn) Fail(prog_.file, "stencil function " + name + " not found");
        kernels_[name] = std::move(ki);
        kernel_order_.push_back(name);
      }
      (reduce ? kernels_[name].used_reduce : kernels_[name].used_plain) = true;
    }
  }
  for (const auto &k : e.kids)
    if (k) FindKernelsIn(*k);
}

void Translator::FindKernelsIn(const Stmt &s) {
  if (s.expr) FindKernelsIn(*s.expr);
  if (s.cond) FindKernelsIn(*s.cond);
  if (s.step) FindKernelsIn(*s.step);
  if (s.init) FindKernelsIn(*s.init);
  for (const auto &d : s.decls)
    if (d.init) FindKernelsIn(*d.init);
  for (const auto &k : s.kids)
    if (k) FindKernelsIn(*k);
}

void Translator::FindKernels() {
  for (auto &it : prog_.items)
    if (it.kind == TopItem::kFunction && it.func->body)
      FindKernelsIn(*it.func->body);
}

bool Translator::RegularIndex(const KernelInfo &k, const Expr &e, int axis,
                              int *off) {
  const Expr *x = &e;
  while (x->kind == EK::kParen) x = x->kids[0].get();
  const std::string &idx = k.index[axis];
  if (x->kind == EK::kIdent && x->text == idx) {
    *off = 0;
    return true;
  }
  if (x->kind == EK::kBinary && (x->op == "+" || x->op == "-")) {
    const Expr *a = x->kids[0].get(), *b = x->kids[1].get();
    while (a->kind == EK::kParen) a = a->kids[0].get();
    int c;
    if (a->kind == EK::kIdent && a->text == idx && ConstInt(*b, &c)) {
      *off = x->op == "+" ? c : -c;
      return true;
    }
    while (b->kind == EK::kParen) b = b->kids[0].get();
    if (x->op == "+" && b->kind == EK::kIdent && b->text == idx &&
        ConstInt(*x->kids[0], &c)) {
      *off = c;
      return true;
    }
  }
  return false;
}

void Translator::AnalyzeGets(KernelInfo &k, const Expr &e) {
  if (const std::string *c = CalleeName(e)) {
    bool get = *c == "PSGridGet", getp = *c == "PSGridGetPeriodic";
    if (get || getp || *c == "PSGridEmit" || *c == "PSGridDim") {
      const Expr *g = e.kids[1].get();
      if (g->kind != EK::kIdent || k.GridIndex(g->text) < 0)
        Fail(k.name, *c + " on something that is not a grid parameter");
      GridParam &gp = k.grids[k.GridIndex(g->text)];
      if (*c == "PSGridEmit") gp.emitted = true;
      if (get || getp) {
        gp.read = true;
        if ((int)e.kids.size() - 2 != gp.dims)
          Fail(k.name, "get of " + gp.name + " with the wrong index count");
        std::array<int, 3> o = {0, 0, 0};
        bool reg = gp.dims == k.ndims;
        for (int a = 0; reg && a < gp.dims; ++a)
          reg = RegularIndex(k, *e.kids[2 + a], a, &o[a]);
        if (!reg) {
          gp.irregular = true;
        } else {
          bool seen = false;
          for (size_t i = 0; i < gp.offsets.size(); ++i)
            if (gp.offsets[i] == o && gp.offset_periodic[i] == getp)
              seen = true;
          if (!seen) {
            gp.offsets.push_back(o);
            gp.offset_periodic.push_back(getp);
          }
        }
      }
    } else if (c->compare(0, 2, "PS") != 0 && c->compare(0, 2, "__") != 0) {
      // a helper the stencil calls: it must be compiled for the device
      for (auto &it : prog_.items)
        if (it.kind == TopItem::kFunction && it.func->name == *c &&
            it.func->body && !device_helpers_.count(*c)) {
          device_helpers_.insert(*c);
          AnalyzeGets(k, *it.func->body);
        }
    }
  }
  for (const auto &kid : e.kids)
    if (kid) AnalyzeGets(k, *kid);
}

void Translator::AnalyzeGets(KernelInfo &k, const Stmt &s) {
  if (s.expr) AnalyzeGets(k, *s.expr);
  if (s.cond) AnalyzeGets(k, *s.cond);
  if (s.step) AnalyzeGets(k, *s.step);
  if (s.init) AnalyzeGets(k, *s.init);
  for (const auto &d : s.decls)
    if (d.init) AnalyzeGets(k, *d.init);
  for (const auto &kid : s.kids)
    if (kid) AnalyzeGets(k, *kid);
}

void Translator::AnalyzeKernel(KernelInfo &k) {
  const Function &f = *k.fn;
  k.ret_type = Strip(f.ret_type);
  size_t i = 0;
  while (i < f.params.size() && IsIntType(f.params[i].type) &&
         !IsGridTypeName(Strip(f.params[i].type), nullptr, nullptr))
    k.index.push_back(f.params[i++].name);
  k.ndims = (int)k.index.size();
  if (k.ndims < 1 || k.ndims > 3)
    Fail(k.name, "a stencil function starts with 1 to 3 index parameters");
  for (; i < f.params.size(); ++i) {
    int nd;
    std::string elem;
    std::string t = Strip(f.params[i].type);
    if (IsGridTypeName(t, &nd, &elem)) {
      GridParam g;
      g.name = f.params[i].name;
      g.dims = nd;
      g.elem = elem;
      g.ps_type = PsTypeOf(elem);
      g.dev_view = "__PSGrid" + std::to_string(nd) + "D" + Cap(elem) + "_dev";
      k.order.push_back({'g', (int)k.grids.size()});
      k.grids.push_back(g);
    } else {
      k.order.push_back({'s', (int)k.scalars.size()});
      k.scalars.push_back({f.params[i].type, f.params[i].name});
    }
  }
  AnalyzeGets(k, *f.body);
  Decide(k);
}

void Translator::Decide(KernelInfo &k) {
  k.fast = k.ndims == 3;
  if (!k.fast) k.slow_reason = "not a 3-D stencil";
  for (auto &g : k.grids)
    if (g.dims != 3 && (g.read || g.emitted)) {
      k.fast = false;
      k.slow_reason = "grid " + g.name + " is not 3-D";
    }
  if (!k.fast) return;
  for (auto &g : k.grids) {
    bool nonzero = false, any_periodic = false;
    for (size_t i = 0; i < g.offsets.size(); ++i) {
      const auto &o = g.offsets[i];
      if (o[0] || o[1] || o[2]) {
        nonzero = true;
        if (g.offset_periodic[i]) any_periodic = true;
      }
    }
    g.periodic = any_periodic;
    if (!g.read || g.offsets.empty()) {
      g.mode = GridParam::kNone;
    } else if (!nonzero) {
      g.mode = GridParam::kDirect;
    } else {
      g.mode = GridParam::kStaged;
      for (size_t i = 0; i < g.offsets.size(); ++i) {
        // non-periodic neighbours of a periodic tile are read from memory
        const auto &o = g.offsets[i];
        if ((o[0] || o[1] || o[2]) && g.offset_periodic[i] != g.periodic)
          continue;
        g.hx = std::max(g.hx, std::abs(o[0]));
        g.hy = std::max(g.hy, std::abs(o[1]));
        g.zl = std::min(g.zl, o[2]);
        g.zr = std::max(g.zr, o[2]);
      }
      if (2 * g.hx > 64) {
        k.fast = false;
        k.slow_reason = "x halo of " + g.name + " wider than a wavefront";
      }
      k.hy = std::max(k.hy, g.hy);
    }
  }
}

// ========================================================= device rewriting
void Translator::RewriteKernelExpr(const KernelInfo &k, ExprP &e) {
  for (auto &kid : e->kids)
    if (kid) RewriteKernelExpr(k, kid);
  const std::string *c = CalleeName(*e);
  if (!c) return;
  bool get = *c == "PSGridGet", getp = *c == "PSGridGetPeriodic";
  if (get || getp) {
    const GridParam &g = *k.Grid(e->kids[1]->text);
    std::array<int, 3> o = {0, 0, 0};
    bool reg = k.fast && g.dims == 3;
    for (int a = 0; reg && a < g.dims; ++a)
      reg = RegularIndex(k, *e->kids[2 + a], a, &o[a]);
    bool zero = o[0] == 0 && o[1] == 0 && o[2] == 0;
    if (reg && g.mode == GridParam::kStaged && (zero || getp == g.periodic)) {
      for (size_t i = 0; i < g.offsets.size(); ++i)
        if (g.offsets[i] == o && (zero || g.offset_periodic[i] == getp)) {
          e = Raw("__ctx." + g.name + "[" + std::to_string(i) + "]");
          return;
        }
    }
    if (reg && g.mode == GridParam::kDirect && zero) {
      e = Raw("__ctx." + g.name + "_c");
      return;
    }
    std::string s = std::string("physis_hip::") +
                    (getp ? "GetPeriodic<" : "GetLinear<") + g.elem +
                    ">(__ctx." + g.name + "_p, __ctx." + g.name + "_dim, " +
                    std::to_string(g.dims);
    for (int a = 0; a < g.dims; ++a) s += ", " + Print(*e->kids[2 + a]);
    e = Raw(s + ")");
    return;
  }
  if (*c == "PSGridEmit") {
    const std::string &n = e->kids[1]->text;
    e = Raw("(__ctx." + n + "_v = (" + Print(*e->kids[2]) + "), __ctx." + n +
            "_e = true)");
    return;
  }
  if (*c == "PSGridDim") {
    const std::string &n = e->kids[1]->text;
    e = Raw("__ctx." + n + "_dim[" + Print(*e->kids[2]) + "]");
    return;
  }
}

void Translator::RewriteKernelStmt(const KernelInfo &k, Stmt &s) {
  if (s.expr) RewriteKernelExpr(k, s.expr);
  if (s.cond) RewriteKernelExpr(k, s.cond);
  if (s.step) RewriteKernelExpr(k, s.step);
  if (s.init) RewriteKernelStmt(k, *s.init);
  for (auto &d : s.decls)
    if (d.init) RewriteKernelExpr(k, d.init);
  for (auto &kid : s.kids)
    if (kid) RewriteKernelStmt(k, *kid);
}

// ============================================================ code emission
std::string Translator::EmitStencilStruct(const KernelInfo &k) {
  std::ostringstream o;
  const std::string &n = k.name;
  o << "struct __PSStencil_" << n << " {\n  __PSDomain dom;\n";
  for (auto &p : k.order) {
    if (p.first == 'g')
      o << "  __PSGrid *" << k.grids[p.second].name << ";\n  int __"
        << k.grids[p.second].name << "_index;\n";
    else
      o << "  " << k.scalars[p.second].type << " "
        << k.scalars[p.second].name << ";\n";
  }
  if (k.ret_type != "void")
    o << "  " << k.ret_type << " *__reduce_out;  // PSStencilMapReduce target\n"
      << "  int __reduce_op;\n";
  o << "};\n\n";
  std::vector<std::string> params = {"__PSDomain dom"}, init = {"dom"};
  for (auto &p : k.order) {
    if (p.first == 'g') {
      const std::string &g = k.grids[p.second].name;
      params.push_back("__PSGrid *" + g);
      init.push_back(g);
      init.push_back("__PSGridGetID(" + g + ")");
    } else {
      params.push_back(k.scalars[p.second].type + " " +
                       k.scalars[p.second].name);
      init.push_back(k.scalars[p.second].name);
    }
  }
  if (k.ret_type != "void") {
    init.push_back("NULL");
    init.push_back("0");
  }
  o << "static struct __PSStencil_" << n << " __PSStencilMap_" << n << "(\n    "
    << Join(params, ", ") << ") {\n  struct __PSStencil_" << n << " s = {"
    << Join(init, ", ") << "};\n  return s;\n}\n\n";
  if (k.ret_type != "void") {
    std::vector<std::string> args;
    for (auto &pp : params) args.push_back(pp.substr(pp.find_last_of(" *") + 1));
    o << "static struct __PSStencil_" << n << " __PSStencilMapReduce_" << n
      << "(\n    " << k.ret_type << " *__v, int __op, " << Join(params, ", ")
      << ") {\n  struct __PSStencil_" << n << " s = __PSStencilMap_" << n << "("
      << Join(args, ", ") << ");\n  s.__reduce_out = __v;\n"
      << "  s.__reduce_op = __op;\n  return s;\n}\n\n";
  }
  return o.str();
}

std::string Translator::EmitContext(const KernelInfo &k) {
  std::ostringstream o;
  o << "// Stencil analysis of " << k.name << " ("
    << (k.fast ? "2.5D run kernel" : "generic run kernel: " + k.slow_reason)
    << "):\n";
  for (auto &g : k.grids) {
    o << "//   " << g.name << "  ";
    if (g.mode == GridParam::kStaged)
      o << "staged x[-" << g.hx << "," << g.hx << "] y[-" << g.hy << ","
        << g.hy << "] z[" << g.zl << "," << g.zr << "]"
        << (g.periodic ? " periodic" : " non-periodic");
    else if (g.mode == GridParam::kDirect)
      o << "direct (centre only)";
    else if (g.read)
      o << "read from memory";
    else
      o << "not read";
    if (g.emitted) o << ", emitted";
    if (g.irregular) o << ", irregular gets";
    o << "\n";
    if (g.mode == GridParam::kStaged) {
      o << "//     offsets";
      for (size_t i = 0; i < g.offsets.size(); ++i)
        o << " " << i << ":(" << g.offsets[i][0] << "," << g.offsets[i][1]
          << "," << g.offsets[i][2] << ")"
          << (g.offset_periodic[i] ? "p" : "");
      o << "\n";
    }
  }
  o << "struct __PSCtx_" << k.name << " {\n";
  for (auto &g : k.grids) {
    if (g.mode == GridParam::kStaged && k.fast)
      o << "  " << g.elem << " " << g.name << "[" << g.offsets.size() << "];\n";
    if (g.mode == GridParam::kDirect && k.fast)
      o << "  " << g.elem << " " << g.name << "_c;\n";
    if (g.emitted)
      o << "  " << g.elem << " " << g.name << "_v;\n  bool " << g.name
        << "_e;\n";
    o << "  const " << g.elem << " *" << g.name << "_p;\n  int " << g.name
      << "_dim[3];\n";
  }
  o << "};\n\n";
  return o.str();
}

std::string Translator::EmitDeviceFunction(KernelInfo &k) {
  Function &f = *const_cast<Function *>(k.fn);
  RewriteKernelStmt(k, *f.body);
  std::ostringstream o;
  std::vector<std::string> ps;
  for (size_t i = 0; i < k.index.size(); ++i)
    ps.push_back(f.params[i].type + " " + f.params[i].name);
  ps.push_back("__PSCtx_" + k.name + " &__ctx");
  for (auto &s : k.scalars) ps.push_back(s.type + " " + s.name);
  o << "__device__ __forceinline__ " << f.ret_type << " " << k.name << "(\n    "
    << Join(ps, ", ") << ") " << Print(*f.body, 0) << "\n";
  return o.str();
}

namespace {
// Distinct (dy, dz) rows of a staged grid with their x extent.
struct Row {
  int dy, dz, xl, xr;
};
std::vector<Row> RowsOf(const GridParam &g) {
  std::vector<Row> rows;
  for (size_t i = 0; i < g.offsets.size(); ++i) {
    const auto &o = g.offsets[i];
    bool zero = !o[0] && !o[1] && !o[2];
    if (!zero && g.offset_periodic[i] != g.periodic) continue;
    bool found = false;
    for (auto &r : rows)
      if (r.dy == o[1] && r.dz == o[2]) {
        r.xl = std::min(r.xl, o[0]);
        r.xr = std::max(r.xr, o[0]);
        found = true;
      }
    if (!found) rows.push_back({o[1], o[2], o[0], o[0]});
  }
  return rows;
}
}  // namespace

std::string Translator::EmitFastRunKernel(const KernelInfo &k) {
  std::ostringstream o;
  const std::string &n = k.name;
  const bool red = k.ret_type != "void";
  const std::string RT = red ? k.ret_type : "float";
  std::vector<std::string> params = {"__PSDomain dom"};
  for (auto &g : k.grids) params.push_back(g.dev_view + " " + g.name);
  for (auto &s : k.scalars) params.push_back(s.type + " " + s.name);
  params.push_back("physis_hip::RunShape sh");
  params.push_back(RT + " *__restrict__ __red");
  o << "// Run kernel: VX points per lane, BY waves per block (the first and\n"
       "// last " << k.hy << " load halo rows), staged planes PD steps ahead,\n"
       "// centre-only grids one step ahead.  RED: reduction operator of a\n"
       "// PSStencilMapReduce execution, or -1.\n";
  o << "template <int VX, int BY, int RED>\n__global__ __launch_bounds__(64 * BY) "
       "void __PSStencilRun_" << n << "(\n    " << Join(params, ", ") << ") {\n";
  o << "  constexpr int PD = " << PrefetchFor(k) << ";\n";
  o << "  constexpr int HY = " << k.hy << ";\n";
  o << "  constexpr int TX = 64 * VX, TY = BY - 2 * HY;\n";
  for (auto &g : k.grids) {
    if (g.mode != GridParam::kStaged) continue;
    int nslot = g.zr - g.zl + 2;
    o << "  typedef physis_hip::Stage<" << g.elem << ", VX, BY, " << g.hx
      << ", HY, " << nslot << ", " << (g.periodic ? "true" : "false") << "> S_"
      << g.name << ";\n";
    o << "  __shared__ __attribute__((aligned(16))) " << g.elem << " "
      << g.name << "_lds[S_" << g.name << "::kLdsElems];\n";
  }
  o << "  const int lane = threadIdx.x % 64;\n"
       "  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);\n"
       "  const physis_hip::TileCoord tc = physis_hip::TileOf(sh.tiles_x, sh.tiles_y);\n"
       "  const int x0 = (sh.x_first_tile + tc.bx) * TX;\n"
       "  const int y0 = dom.local_min[1] + tc.by * TY;\n"
       "  const int zb = dom.local_min[2] + tc.bz * sh.zchunk;\n"
       "  const int ze = min(zb + sh.zchunk, dom.local_max[2]);\n";
  if (red) o << "  " << RT << " __racc = 0;\n"
             << "  if constexpr (RED >= 0) __racc = physis_hip::ReduceOp<" << RT
             << ", RED < 0 ? 0 : RED>::identity();\n";
  o << "  if (zb < ze) {\n"
       "    const int yrow = y0 - HY + wave;\n"
       "    const bool computes = wave >= HY && wave < BY - HY && yrow < dom.local_max[1];\n"
       "    const int xs = x0 + lane * VX;\n";
  bool any_direct = false;
  for (auto &g : k.grids) any_direct = any_direct || g.mode == GridParam::kDirect;
  if (any_direct)
    o << "    // centre-only grids are read at the nearest compute row, so halo\n"
         "    // waves repeat their neighbours' addresses and every wave issues\n"
         "    // the same loads\n"
         "    const int drow = wave < HY ? y0 : (wave >= BY - HY ? y0 + TY - 1 : yrow);\n";
  for (auto &g : k.grids) {
    const std::string &G = g.name;
    if (g.mode == GridParam::kStaged) {
      int nslot = g.zr - g.zl + 2;
      o << "    const int " << G << "_nxy = " << G << ".dim[0] * " << G << ".dim[1];\n";
      o << "    const physis_hip::BufWin<" << g.elem << "> " << G
        << "_w = physis_hip::BufWin<" << g.elem << ">::Make(" << G << ".p, "
        << G << "_nxy * " << G << ".dim[2], (long)(zb + " << g.zl - 1 << ") * "
        << G << "_nxy, (long)(ze + " << g.zr + 1 << ") * " << G << "_nxy);\n";
      o << "    S_" << G << " " << G << "_st[PD];\n";
      o << "#pragma unroll\n    for (int k = 0; k < " << nslot - 1 << "; ++k) {\n"
        << "      " << G << "_st[0].Load(" << G << "_w, " << G << ".p, " << G
        << ".dim, x0, yrow, zb + " << g.zl << " + k, lane);\n"
        << "      " << G << "_st[0].Store(" << G << "_lds + k * S_" << G
        << "::kPlane, wave, lane);\n    }\n";
      o << "#pragma unroll\n    for (int k = 0; k < PD; ++k)\n      " << G
        << "_st[k].Load(" << G << "_w, " << G << ".p, " << G
        << ".dim, x0, yrow, zb + " << g.zr + 1 << " + k, lane);\n";
      o << "    int " << G << "_base = 0;\n";
    } else if (g.mode == GridParam::kDirect) {
      o << "    const int " << G << "_nx = " << G << ".dim[0], " << G
        << "_nxy = " << G << ".dim[0] * " << G << ".dim[1];\n";
      o << "    const physis_hip::BufWin<" << g.elem << "> " << G
        << "_w = physis_hip::BufWin<" << g.elem << ">::Make(" << G << ".p, "
        << G << "_nxy * " << G << ".dim[2], (long)zb * " << G
        << "_nxy, (long)ze * " << G << "_nxy);\n";
      o << "    " << g.elem << " " << G << "_dv[2][VX];\n";
      o << "    " << G << "_w.template Load<VX>(xs + drow * " << G
        << "_nx + zb * " << G << "_nxy, " << G << "_dv[0]);\n";
    }
    if (g.emitted) {
      o << "    const int " << G << "_enx = " << G << ".dim[0], " << G
        << "_enxy = " << G << ".dim[0] * " << G << ".dim[1];\n";
      o << "    const physis_hip::BufWin<" << g.elem << "> " << G
        << "_ew = physis_hip::BufWin<" << g.elem << ">::Make(" << G << ".p, "
        << G << "_enxy * " << G << ".dim[2], (long)zb * " << G
        << "_enxy, (long)ze * " << G << "_enxy);\n";
    }
  }
  o << "    __syncthreads();\n";
  o << "    __PSCtx_" << n << " ctx;\n";
  for (auto &g : k.grids)
    o << "    ctx." << g.name << "_p = " << g.name << ".p;\n"
      << "    for (int d = 0; d < 3; ++d) ctx." << g.name << "_dim[d] = "
      << g.name << ".dim[d];\n";
  o << "    for (int zc = zb; zc < ze; zc += PD) {\n#pragma unroll\n"
       "      for (int k = 0; k < PD; ++k) {\n"
       "        const int z = zc + k;\n"
       "        const int cur = k & 1, nxt = cur ^ 1;\n"
       "        (void)cur; (void)nxt;\n";
  for (auto &g : k.grids)
    if (g.mode == GridParam::kDirect)
      o << "        " << g.name << "_w.template Load<VX>(xs + drow * " << g.name
        << "_nx + (z + 1) * " << g.name << "_nxy, " << g.name << "_dv[nxt]);\n";
  o << "        if (computes && z < ze) {\n";
  for (auto &g : k.grids) {
    if (g.mode != GridParam::kStaged) continue;
    int nslot = g.zr - g.zl + 2;
    auto rows = RowsOf(g);
    for (size_t r = 0; r < rows.size(); ++r) {
      const Row &w = rows[r];
      o << "          " << g.elem << " " << g.name << "_r" << r << "[VX + "
        << w.xr - w.xl << "];\n";
      o << "          S_" << g.name << "::template ReadRow<" << w.xl << ", "
        << w.xr << ">(" << g.name << "_lds + physis_hip::RingSlot<" << nslot
        << ">(" << g.name << "_base, " << w.dz - g.zl << ") * S_" << g.name
        << "::kPlane, wave + " << w.dy << ", lane, " << g.name << "_r" << r
        << ");\n";
    }
  }
  for (auto &g : k.grids)
    if (g.emitted)
      o << "          " << g.elem << " " << g.name << "_out[VX];\n"
        << "          unsigned " << g.name << "_mask = 0;\n";
  o << "#pragma unroll\n          for (int i = 0; i < VX; ++i) {\n"
       "            const int x = xs + i;\n";
  for (auto &g : k.grids)
    if (g.emitted) o << "            " << g.name << "_out[i] = 0;\n";
  o << "            if (x >= dom.local_min[0] && x < dom.local_max[0]) {\n";
  for (auto &g : k.grids) {
    if (g.mode == GridParam::kStaged) {
      auto rows = RowsOf(g);
      for (size_t i = 0; i < g.offsets.size(); ++i) {
        const auto &off = g.offsets[i];
        bool zero = !off[0] && !off[1] && !off[2];
        if (!zero && g.offset_periodic[i] != g.periodic) continue;
        for (size_t r = 0; r < rows.size(); ++r)
          if (rows[r].dy == off[1] && rows[r].dz == off[2])
            o << "              ctx." << g.name << "[" << i << "] = " << g.name
              << "_r" << r << "[i + " << off[0] - rows[r].xl << "];\n";
      }
    } else if (g.mode == GridParam::kDirect) {
      o << "              ctx." << g.name << "_c = " << g.name << "_dv[cur][i];\n";
    }
    if (g.emitted) o << "              ctx." << g.name << "_e = false;\n";
  }
  std::vector<std::string> args = {"x", "yrow", "z", "ctx"};
  for (auto &s : k.scalars) args.push_back(s.name);
  if (red)
    o << "              const " << RT << " __r = " << n << "(" << Join(args, ", ")
      << ");\n              if constexpr (RED >= 0) __racc = physis_hip::ReduceOp<"
      << RT << ", RED < 0 ? 0 : RED>::apply(__racc, __r);\n";
  else
    o << "              " << n << "(" << Join(args, ", ") << ");\n";
  for (auto &g : k.grids)
    if (g.emitted)
      o << "              if (ctx." << g.name << "_e) {\n                "
        << g.name << "_out[i] = ctx." << g.name << "_v;\n                "
        << g.name << "_mask |= 1u << i;\n              }\n";
  o << "            }\n          }\n";
  for (auto &g : k.grids)
    if (g.emitted)
      o << "          physis_hip::StoreMasked<" << g.elem << ", VX>(" << g.name
        << "_ew, xs + yrow * " << g.name << "_enx + z * " << g.name
        << "_enxy, " << g.name << "_out, " << g.name << "_mask);\n";
  o << "        }\n";
  for (auto &g : k.grids) {
    if (g.mode != GridParam::kStaged) continue;
    int nslot = g.zr - g.zl + 2;
    o << "        " << g.name << "_st[k].Store(" << g.name
      << "_lds + physis_hip::RingSlot<" << nslot << ">(" << g.name << "_base, "
      << nslot - 1 << ") * S_" << g.name << "::kPlane, wave, lane);\n";
  }
  o << "        __syncthreads();\n";
  for (auto &g : k.grids) {
    if (g.mode != GridParam::kStaged) continue;
    int nslot = g.zr - g.zl + 2;
    o << "        " << g.name << "_st[k].Load(" << g.name << "_w, " << g.name
      << ".p, " << g.name << ".dim, x0, yrow, z + " << g.zr + 1
      << " + PD, lane);\n"
      << "        " << g.name << "_base = physis_hip::RingSlot<" << nslot
      << ">(" << g.name << "_base, 1);\n";
  }
  o << "      }\n    }\n  }\n";
  if (red)
    o << "  if constexpr (RED >= 0) {\n    __shared__ " << RT
      << " __lds_red[BY];\n    const " << RT
      << " __r = physis_hip::BlockReduce<" << RT
      << ", RED < 0 ? 0 : RED, 64 * BY>(__racc, __lds_red);\n"
         "    if (threadIdx.x == 0) __red[blockIdx.x] = __r;\n  }\n";
  else
    o << "  (void)__red;\n";
  o << "}\n\n";
  return o.str();
}

std::string Translator::EmitGenericRunKernel(const KernelInfo &k) {
  std::ostringstream o;
  const std::string &n = k.name;
  const bool red = k.ret_type != "void";
  const std::string RT = red ? k.ret_type : "float";
  std::vector<std::string> params = {"__PSDomain dom"};
  for (auto &g : k.grids) params.push_back(g.dev_view + " " + g.name);
  for (auto &s : k.scalars) params.push_back(s.type + " " + s.name);
  params.push_back(RT + " *__restrict__ __red");
  o << "// Generic run kernel (one thread per (x, y) column marching z, all\n"
       "// gets read from memory): " << k.slow_reason << ".\n";
  o << "template <int RED>\n__global__ __launch_bounds__(256) void "
       "__PSStencilRunGeneric_" << n << "(\n    " << Join(params, ", ")
    << ") {\n";
  o << "  const int x = dom.local_min[0] + blockIdx.x * 64 + threadIdx.x % 64;\n";
  if (k.ndims >= 2)
    o << "  const int y = dom.local_min[1] + blockIdx.y * 4 + threadIdx.x / 64;\n";
  if (red)
    o << "  " << RT << " __racc = 0;\n  if constexpr (RED >= 0) __racc = "
         "physis_hip::ReduceOp<" << RT << ", RED < 0 ? 0 : RED>::identity();\n";
  o << "  if (x < dom.local_max[0]" << (k.ndims >= 2 ? " && y < dom.local_max[1]" : "")
    << ") {\n    __PSCtx_" << n << " ctx;\n";
  for (auto &g : k.grids)
    o << "    ctx." << g.name << "_p = " << g.name << ".p;\n"
      << "    for (int d = 0; d < 3; ++d) ctx." << g.name << "_dim[d] = "
      << (g.dims == 1 ? "d < 1 ? " + g.name + ".dim[d] : 1"
                      : g.dims == 2 ? "d < 2 ? " + g.name + ".dim[d] : 1"
                                    : g.name + ".dim[d]")
      << ";\n";
  if (k.ndims == 3)
    o << "    for (int z = dom.local_min[2]; z < dom.local_max[2]; ++z) {\n";
  else
    o << "    {\n";
  for (auto &g : k.grids)
    if (g.emitted) o << "      ctx." << g.name << "_e = false;\n";
  std::vector<std::string> args = {"x"};
  if (k.ndims >= 2) args.push_back("y");
  if (k.ndims >= 3) args.push_back("z");
  args.push_back("ctx");
  for (auto &s : k.scalars) args.push_back(s.name);
  if (red)
    o << "      const " << RT << " __r = " << n << "(" << Join(args, ", ")
      << ");\n      if constexpr (RED >= 0) __racc = physis_hip::ReduceOp<" << RT
      << ", RED < 0 ? 0 : RED>::apply(__racc, __r);\n";
  else
    o << "      " << n << "(" << Join(args, ", ") << ");\n";
  for (auto &g : k.grids) {
    if (!g.emitted) continue;
    std::string off = "x";
    if (g.dims >= 2) off += " + (long)y * " + g.name + ".dim[0]";
    if (g.dims >= 3)
      off += " + (long)z * " + g.name + ".dim[0] * " + g.name + ".dim[1]";
    o << "      if (ctx." << g.name << "_e) " << g.name << ".p[" << off
      << "] = ctx." << g.name << "_v;\n";
  }
  o << "    }\n  }\n";
  if (red)
    o << "  if constexpr (RED >= 0) {\n    __shared__ " << RT
      << " __lds_red[4];\n    const " << RT << " __r = physis_hip::BlockReduce<"
      << RT << ", RED < 0 ? 0 : RED, 256>(__racc, __lds_red);\n"
         "    if (threadIdx.x == 0) __red[blockIdx.y * gridDim.x + blockIdx.x] = __r;\n  }\n";
  else
    o << "  (void)__red;\n";
  o << "}\n\n";
  return o.str();
}

int Translator::RegEstimate(const KernelInfo &k, int pd) const {
  // double-buffered centre values, the LDS row segments, the staged planes
  // in flight, addressing
  int est = 48;
  for (auto &g : k.grids) {
    int w = (g.elem == "double" || g.elem == "long") ? 2 : 1;
    if (g.mode == GridParam::kDirect) est += 2 * 4 * w;
    if (g.mode == GridParam::kStaged) {
      est += pd * 5 * w;
      for (auto &r : RowsOf(g)) est += (4 + r.xr - r.xl) * w;
    }
  }
  return est;
}

// Four staged planes in flight hide HBM latency better than two when the
// kernel stays well inside 128 VGPRs (measured on the 512^3 7-point kernel:
// 198 us against 211 us per sweep); kernels with many centre-only grids
// keep two (the 19-point Himeno kernel would exceed the register file).
int Translator::PrefetchFor(const KernelInfo &k) const {
  if (cfg_.prefetch > 0) return cfg_.prefetch;
  return RegEstimate(k, 4) <= 96 ? 4 : 2;
}

std::string Translator::EmitLauncher(const KernelInfo &k) {
  std::ostringstream o;
  const std::string &n = k.name;
  const bool red = k.ret_type != "void";
  const std::string RT = red ? k.ret_type : "float";
  std::vector<std::string> args = {"s.dom"};
  for (auto &g : k.grids)
    args.push_back("*(" + g.dev_view + " *)s." + g.name + "->dev");
  for (auto &s : k.scalars) args.push_back("s." + s.name);
  std::string A = Join(args, ", ");
  if (k.fast) {
    // LDS budget picks the block height.
    int by = cfg_.block_waves;
    if (by == 0) {
      // A 1024-thread block allows 128 VGPRs per lane, a 512-thread one 256.
      const int est = RegEstimate(k, PrefetchFor(k));
      by = 4;
      for (int cand : {16, 8, 4}) {
        if (cand - 2 * k.hy < 1) continue;
        if (cand == 16 && est > 128) continue;
        long lds = 0;
        for (auto &g : k.grids)
          if (g.mode == GridParam::kStaged) {
            int elm = (g.elem == "double" || g.elem == "long") ? 8 : 4;
            int pad = g.hx == 0 ? 0 : ((g.hx + 3) / 4) * 4;
            lds += (long)(g.zr - g.zl + 2) * cand * (2 * pad + 256) * elm;
          }
        if (lds <= 80 * 1024) {
          by = cand;
          break;
        }
      }
    }
    if (by - 2 * k.hy < 1) Fail(n, "y halo too deep for the run kernel");
    o << "template <int VX>\nstatic physis_hip::RunShape __PSShape_" << n
      << "(const struct __PSStencil_" << n << " &s) {\n"
      << "  static int per_cu = physis_hip::BlocksPerCU(\n"
      << "      (const void *)__PSStencilRun_" << n << "<VX, " << by
      << ", -1>, 64 * " << by << ");\n"
      << "  return physis_hip::PlanRun(s.dom, 64 * VX, " << by << " - 2 * "
      << k.hy << ", per_cu);\n}\n\n";
    o << "template <int VX>\nstatic int __PSLaunch_" << n << "(const struct __PSStencil_"
      << n << " &s, hipStream_t stream, void *red) {\n"
      << "  const physis_hip::RunShape sh = __PSShape_" << n << "<VX>(s);\n"
      << "  if (sh.blocks == 0) return 0;\n"
      << "  const dim3 grid(sh.blocks), block(64 * " << by << ");\n";
    if (red && k.used_reduce) {
      o << "  switch (red ? s.__reduce_op : -1) {\n";
      for (const char *op : {"PS_MAX", "PS_MIN", "PS_SUM", "PS_PROD"})
        o << "    case " << op << ":\n      hipLaunchKernelGGL((__PSStencilRun_" << n
          << "<VX, " << by << ", " << op << ">), grid, block, 0, stream, " << A
          << ", sh, (" << RT << " *)red);\n      break;\n";
      o << "    default:\n";
    } else {
      o << "  (void)red;\n  {\n";
    }
    o << "      hipLaunchKernelGGL((__PSStencilRun_" << n << "<VX, " << by
      << ", -1>), grid, block, 0, stream, " << A << ", sh, (" << RT
      << " *)NULL);\n  }\n  return sh.blocks;\n}\n\n";
    // vector width: every 3-D grid's x extent a multiple of 4
    std::vector<std::string> conds;
    for (auto &g : k.grids)
      if (g.read || g.emitted) conds.push_back("s." + g.name + "->dim[0] % 4 == 0");
    std::string vec = conds.empty() ? "true" : Join(conds, " && ");
    o << "static int __PSBlocks_" << n << "(const struct __PSStencil_" << n
      << " &s) {\n  return (" << vec << ") ? __PSShape_" << n
      << "<4>(s).blocks : __PSShape_" << n << "<1>(s).blocks;\n}\n\n";
    o << "static int __PSRun_" << n << "(const struct __PSStencil_" << n
      << " &s, hipStream_t stream, void *red) {\n";
    for (auto &g : k.grids)
      if (g.read || g.emitted)
        o << "  physis_hip::CheckGrid3D(s." << g.name << ", " << g.ps_type << ");\n";
    for (auto &g : k.grids)
      if (g.emitted)
        o << "  physis_hip::CheckEmitDomain(s.dom, s." << g.name << ");\n";
    o << "  return (" << vec << ") ? __PSLaunch_" << n
      << "<4>(s, stream, red) : __PSLaunch_" << n << "<1>(s, stream, red);\n}\n\n";
  } else {
    o << "static dim3 __PSGridDim_" << n << "(const struct __PSStencil_" << n
      << " &s) {\n  const int nx = s.dom.local_max[0] - s.dom.local_min[0];\n"
      << "  const int ny = " << (k.ndims >= 2 ? "s.dom.local_max[1] - s.dom.local_min[1]" : "1")
      << ";\n  if (nx <= 0 || ny <= 0) return dim3(0, 0, 1);\n"
      << "  return dim3((nx + 63) / 64, " << (k.ndims >= 2 ? "(ny + 3) / 4" : "1")
      << ", 1);\n}\n\n";
    o << "static int __PSBlocks_" << n << "(const struct __PSStencil_" << n
      << " &s) {\n  const dim3 g = __PSGridDim_" << n
      << "(s);\n  return g.x * g.y;\n}\n\n";
    o << "static int __PSRun_" << n << "(const struct __PSStencil_" << n
      << " &s, hipStream_t stream, void *red) {\n";
    for (auto &g : k.grids)
      if (g.emitted)
        o << "  physis_hip::CheckEmitDomain(s.dom, s." << g.name << ");\n";
    if (k.ndims == 3)
      o << "  if (s.dom.local_max[2] <= s.dom.local_min[2]) return 0;\n";
    o << "  const dim3 grid = __PSGridDim_" << n << "(s);\n"
      << "  if (grid.x == 0) return 0;\n";
    if (red && k.used_reduce) {
      o << "  switch (red ? s.__reduce_op : -1) {\n";
      for (const char *op : {"PS_MAX", "PS_MIN", "PS_SUM", "PS_PROD"})
        o << "    case " << op << ":\n      hipLaunchKernelGGL((__PSStencilRunGeneric_"
          << n << "<" << op << ">), grid, dim3(256), 0, stream, " << A << ", ("
          << RT << " *)red);\n      break;\n";
      o << "    default:\n";
    } else {
      o << "  (void)red;\n  {\n";
    }
    o << "      hipLaunchKernelGGL((__PSStencilRunGeneric_" << n
      << "<-1>), grid, dim3(256), 0, stream, " << A << ", (" << RT
      << " *)NULL);\n  }\n  return grid.x * grid.y;\n}\n\n";
  }
  return o.str();
}

std::string Translator::EmitKernel(KernelInfo &k) {
  std::string s = "// ---------------------------------------------- stencil " +
                  k.name + "\n";
  s += EmitStencilStruct(k);
  s += EmitContext(k);
  s += EmitDeviceFunction(k);
  s += k.fast ? EmitFastRunKernel(k) : EmitGenericRunKernel(k);
  s += EmitLauncher(k);
  return s;
}

std::string Translator::EmitRunFunction(const RunSite &r) {
  std::ostringstream o;
  std::vector<std::string> params = {"int iter"};
  for (size_t i = 0; i < r.kernels.size(); ++i)
    params.push_back("struct __PSStencil_" + r.kernels[i] + " s" +
                     std::to_string(i));
  o << "static void __PSStencilRun_" << r.id << "(" << Join(params, ", ")
    << ") {\n  hipStream_t stream = (hipStream_t)__PSHipStream();\n"
    << "  __PSTraceStencilPre(\"" << Join(r.kernels, ", ") << "\");\n"
    << "  physis_hip::RunTimer timer(stream);\n";
  for (size_t i = 0; i < r.kernels.size(); ++i) {
    const KernelInfo &k = kernels_.at(r.kernels[i]);
    std::string si = "s" + std::to_string(i);
    if (k.ret_type != "void")
      o << "  physis_hip::MapReducer red" << i << "(" << si << ".__reduce_out, "
        << si << ".__reduce_op, " << PsTypeOf(k.ret_type) << ",\n      " << si
        << ".__reduce_out ? __PSBlocks_" << k.name << "(" << si
        << ") : 0, iter, " << i << ", " << r.kernels.size() << ");\n";
  }
  o << "  for (int i = 0; i < iter; ++i) {\n";
  for (size_t i = 0; i < r.kernels.size(); ++i) {
    const KernelInfo &k = kernels_.at(r.kernels[i]);
    o << "    __PSRun_" << k.name << "(s" << i << ", stream, "
      << (k.ret_type != "void" ? "red" + std::to_string(i) + ".Slot(i)" : "NULL")
      << ");\n";
  }
  for (size_t i = 0; i < r.kernels.size(); ++i)
    if (kernels_.at(r.kernels[i]).ret_type != "void")
      o << "    red" << i << ".After(i);\n";
  o << "  }\n";
  if (cfg_.kernel_error_check) o << "  __PSCheckHipError(\"Kernel Execution Failed!\");\n";
  o << "  __PSTraceStencilPost(timer.Stop());\n}\n\n";
  return o.str();
}

// ============================================================= host code
std::string Translator::GridElemOf(const Expr &g, const Function &f) const {
  const Expr *x = &g;
  while (x->kind == EK::kParen) x = x->kids[0].get();
  if (x->kind != EK::kIdent) return "";
  std::string type;
  auto lit = local_grids_.find(f.name);
  if (lit != local_grids_.end() && lit->second.count(x->text))
    type = lit->second.at(x->text);
  else if (grid_vars_.count(x->text))
    type = grid_vars_.at(x->text);
  std::string elem;
  if (!IsGridTypeName(type, nullptr, &elem)) return "";
  return elem;
}

ExprP Translator::MapCall(Expr &call, const Function &f) {
  const std::string &c = *CalleeName(call);
  (void)f;
  auto out = std::make_unique<Expr>(EK::kCall);
  auto callee = std::make_unique<Expr>(EK::kIdent);
  if (c == "PSStencilMap") {
    std::string k = call.kids[1]->text;
    while (call.kids[1]->kind == EK::kParen || call.kids[1]->kind == EK::kCast)
      call.kids[1] = std::move(call.kids[1]->kids[0]);
    callee->text = "__PSStencilMap_" + call.kids[1]->text;
    out->kids.push_back(std::move(callee));
    for (size_t i = 2; i < call.kids.size(); ++i)
      out->kids.push_back(std::move(call.kids[i]));
  } else {  // PSStencilMapReduce(v, op, k, dom, ...)
    while (call.kids[3]->kind == EK::kParen || call.kids[3]->kind == EK::kCast)
      call.kids[3] = std::move(call.kids[3]->kids[0]);
    callee->text = "__PSStencilMapReduce_" + call.kids[3]->text;
    out->kids.push_back(std::move(callee));
    out->kids.push_back(std::move(call.kids[1]));
    out->kids.push_back(std::move(call.kids[2]));
    for (size_t i = 4; i < call.kids.size(); ++i)
      out->kids.push_back(std::move(call.kids[i]));
  }
  return out;
}

void Translator::RewriteHostExpr(ExprP &e, const Function &f) {
  const std::string *cp = CalleeName(*e);
  std::string c = cp ? *cp : "";
  if (c == "PSStencilRun") {
    RunSite site;
    site.id = (int)runs_.size();
    auto call = std::make_unique<Expr>(EK::kCall);
    call->kids.push_back(Raw("__PSStencilRun_" + std::to_string(site.id)));
    ExprP iter = Raw("1");
    std::vector<ExprP> maps;
    for (size_t i = 1; i < e->kids.size(); ++i) {
      Expr &a = *e->kids[i];
      const std::string *ac = CalleeName(a);
      if (ac && (*ac == "PSStencilMap" || *ac == "PSStencilMapReduce")) {
        const Expr *kn = (*ac == "PSStencilMap" ? a.kids[1] : a.kids[3]).get();
        while (kn->kind == EK::kParen || kn->kind == EK::kCast) kn = kn->kids[0].get();
        site.kernels.push_back(kn->text);
        for (size_t j = 2; j < a.kids.size(); ++j) RewriteHostExpr(a.kids[j], f);
        maps.push_back(MapCall(a, f));
      } else if (i + 1 == e->kids.size()) {
        RewriteHostExpr(e->kids[i], f);
        iter = std::move(e->kids[i]);
      } else {
        Fail(f.name, "PSStencilRun arguments must be PSStencilMap calls");
      }
    }
    call->kids.push_back(std::move(iter));
    for (auto &m : maps) call->kids.push_back(std::move(m));
    runs_in_[&f].push_back(site.id);
    runs_.push_back(site);
    e = std::move(call);
    return;
  }
  if (c == "PSReduce" && e->kids.size() >= 5) {  // kernel form
    RunSite site;
    site.id = (int)runs_.size();
    for (size_t j = 1; j < e->kids.size(); ++j) RewriteHostExpr(e->kids[j], f);
    auto m = std::make_unique<Expr>(EK::kCall);
    m->kids.push_back(Raw("PSStencilMapReduce"));
    for (size_t j = 1; j < e->kids.size(); ++j)
      m->kids.push_back(std::move(e->kids[j]));
    const Expr *kn = m->kids[3].get();
    while (kn->kind == EK::kParen || kn->kind == EK::kCast) kn = kn->kids[0].get();
    site.kernels.push_back(kn->text);
    const KernelInfo &k = kernels_.at(kn->text);
    // the reduced value lands in the caller's variable
    ExprP v = std::move(m->kids[1]);
    m->kids[1] = Raw("(" + k.ret_type + " *)(" + Print(*v) + ")");
    auto call = std::make_unique<Expr>(EK::kCall);
    call->kids.push_back(Raw("__PSStencilRun_" + std::to_string(site.id)));
    call->kids.push_back(Raw("1"));
    call->kids.push_back(MapCall(*m, f));
    runs_in_[&f].push_back(site.id);
    runs_.push_back(site);
    e = std::move(call);
    return;
  }
  for (auto &kid : e->kids)
    if (kid) RewriteHostExpr(kid, f);
  if (c.empty()) return;
  int nd;
  std::string elem;
  if (c.size() > 3 && c.compare(c.size() - 3, 3, "New") == 0 &&
      IsGridTypeName(c.substr(0, c.size() - 3), &nd, &elem)) {
    std::string s = "physis_hip::NewGrid(" + PsTypeOf(elem) + ", sizeof(" +
                    elem + "), " + std::to_string(nd);
    for (int a = 0; a < nd; ++a) s += ", " + Print(*e->kids[1 + a]);
    e = Raw(s + ")");
  } else if (c == "PSGridCopyin" || c == "PSGridCopyout") {
    e = Raw("__" + c + "(" + Print(*e->kids[1]) + ", " + Print(*e->kids[2]) +
            ", NULL)");
  } else if (c == "PSGridFree") {
    e = Raw("__PSGridFree(" + Print(*e->kids[1]) + ", NULL)");
  } else if (c == "PSGridSet") {
    std::string el = GridElemOf(*e->kids[1], f);
    if (el.empty()) Fail(f.name, "PSGridSet on a grid of unknown type");
    std::string s = "physis_hip::SetElem<" + el + ">(" + Print(*e->kids[1]) +
                    ", " + Print(*e->kids.back());
    for (size_t a = 2; a + 1 < e->kids.size(); ++a) s += ", " + Print(*e->kids[a]);
    e = Raw(s + ")");
  } else if (c == "PSReduce") {
    std::string el = GridElemOf(*e->kids[3], f);
    if (el.empty()) Fail(f.name, "PSReduce on a grid of unknown type");
    e = Raw(ReduceFnOf(el) + "(" + Print(*e->kids[1]) + ", (enum PSReduceOp)(" +
            Print(*e->kids[2]) + "), " + Print(*e->kids[3]) + ")");
  }
}

void Translator::RewriteHostStmt(Stmt &s, const Function &f) {
  if (s.kind == SK::kDecl) {
    std::istringstream is(s.type);
    std::string w;
    while (is >> w)
      if (IsGridTypeName(w, nullptr, nullptr))
        for (auto &d : s.decls) local_grids_[f.name][d.name] = w;
  }
  if (s.expr) RewriteHostExpr(s.expr, f);
  if (s.cond) RewriteHostExpr(s.cond, f);
  if (s.step) RewriteHostExpr(s.step, f);
  if (s.init) RewriteHostStmt(*s.init, f);
  for (auto &d : s.decls)
    if (d.init) RewriteHostExpr(d.init, f);
  for (auto &k : s.kids)
    if (k) RewriteHostStmt(*k, f);
}

// ================================================================= driver
std::string Translator::Translate() {
  FindKernels();
  for (auto &n : kernel_order_) AnalyzeKernel(kernels_[n]);

  std::ostringstream out;
  out << "// Generated by physisc-hip from " << prog_.file << ".\n"
      << "#define PHYSIS_HIP\n#include \"physis/physis.h\"\n"
      << "#include \"physis/hip/stencil.h\"\n#include \"physis/hip/run.h\"\n";
  for (auto &it : prog_.items)
    if (it.kind == TopItem::kInclude) out << it.include_line << "\n";
  out << "\n#pragma clang fp contract(off)\n\n";

  // host functions: rewrite first so that run sites are known
  for (auto &it : prog_.items) {
    if (it.kind != TopItem::kFunction) continue;
    Function &f = *it.func;
    if (kernels_.count(f.name)) continue;
    for (auto &p : f.params) {
      std::string t = Strip(p.type);
      if (IsGridTypeName(t, nullptr, nullptr)) local_grids_[f.name][p.name] = t;
    }
    if (f.body) RewriteHostStmt(*f.body, f);
  }

  bool in_c = false;
  auto open_c = [&]() {
    if (!in_c) out << "extern \"C\" {\n\n";
    in_c = true;
  };
  auto close_c = [&]() {
    if (in_c) out << "}  // extern \"C\"\n\n";
    in_c = false;
  };
  std::set<std::string> emitted_kernels;
  for (auto &it : prog_.items) {
    if (it.kind == TopItem::kInclude) continue;
    if (it.kind == TopItem::kTokens) {
      open_c();
      out << PrintTokens(it.tokens) << "\n\n";
      continue;
    }
    Function &f = *it.func;
    if (kernels_.count(f.name)) {
      close_c();
      out << EmitKernel(kernels_[f.name]);
      emitted_kernels.insert(f.name);
      continue;
    }
    std::vector<std::string> ps;
    for (auto &p : f.params) ps.push_back(p.type + " " + p.name);
    if (f.variadic) ps.push_back("...");
    if (ps.empty()) ps.push_back("void");
    if (device_helpers_.count(f.name)) {
      close_c();
      out << (f.is_static ? "static " : "") << "__host__ __device__ "
          << (f.is_inline ? "inline " : "") << f.ret_type << " " << f.name
          << "(" << Join(ps, ", ") << ") " << Print(*f.body, 0) << "\n";
      continue;
    }
    auto rit = runs_in_.find(&f);
    if (rit != runs_in_.end()) {
      close_c();
      for (int id : rit->second) {
        for (auto &kn : runs_[id].kernels)
          if (!emitted_kernels.count(kn))
            Fail(f.name, "stencil " + kn + " must be defined before its use");
        out << EmitRunFunction(runs_[id]);
      }
    }
    open_c();
    out << (f.is_static ? "static " : "") << (f.is_inline ? "inline " : "")
        << f.ret_type << " " << f.name << "(" << Join(ps, ", ") << ") "
        << Print(*f.body, 0) << "\n";
  }
  close_c();
  return out.str();
}

}  // namespace physisc
