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

// PSStencilMap and its red-black forms (physis_user.h:98-105); the colour
// argument of the run kernel: -1 all points, 0 red, 1 black, 2 red then
// black.
bool IsMapName(const std::string &c, int *rb = nullptr) {
  int v = -2;
  if (c == "PSStencilMap") v = -1;
  else if (c == "PSStencilMapRed") v = 0;
  else if (c == "PSStencilMapBlack") v = 1;
  else if (c == "PSStencilMapRedBlack") v = 2;
  if (rb) *rb = v;
  return v != -2;
}

bool MentionsIdent(const Expr &e, const std::string &name) {
  if (e.kind == EK::kIdent && e.text == name) return true;
  for (const auto &k : e.kids)
    if (k && MentionsIdent(*k, name)) return true;
  return false;
}

// Whether a comparison (== != < > <= >=) in e mentions `name`.
bool ComparesIdent(const Expr &e, const std::string &name) {
  if (e.kind == EK::kBinary &&
      (e.op == "==" || e.op == "!=" || e.op == "<" || e.op == ">" ||
       e.op == "<=" || e.op == ">=") &&
      MentionsIdent(e, name))
    return true;
  for (const auto &k : e.kids)
    if (k && ComparesIdent(*k, name)) return true;
  return false;
}

bool ComparesIdent(const Stmt &s, const std::string &name) {
  if (s.expr && ComparesIdent(*s.expr, name)) return true;
  if (s.cond && ComparesIdent(*s.cond, name)) return true;
  if (s.step && ComparesIdent(*s.step, name)) return true;
  if (s.init && ComparesIdent(*s.init, name)) return true;
  for (const auto &d : s.decls)
    if (d.init && ComparesIdent(*d.init, name)) return true;
  for (const auto &k : s.kids)
    if (k && ComparesIdent(*k, name)) return true;
  return false;
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
    int rb = -2;
    if (IsMapName(*c, &rb) && e.kids.size() >= 3) karg = 1;
    if (*c == "PSStencilMapReduce" && e.kids.size() >= 5) {
      karg = 3;
      reduce = true;
    }
    if (*c == "PSReduce" && e.kids.size() >= 5) {  // kernel form
      karg = 3;
      reduce = true;
    }
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
      if (rb >= 0) kernels_[name].used_rb = true;
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

// A function shaped like a stencil (index parameters, then grid parameters)
// that emits, but that no map uses: its body only has meaning on the device,
// so it is left out rather than compiled as host code.
bool Translator::IsUnusedStencil(const Function &f) const {
  if (!f.body || kernels_.count(f.name) || device_helpers_.count(f.name))
    return false;
  size_t i = 0;
  while (i < f.params.size() && IsIntType(f.params[i].type)) ++i;
  if (i == 0 || i > 3 || i == f.params.size() ||
      !IsGridTypeName(Strip(f.params[i].type), nullptr, nullptr))
    return false;
  std::function<bool(const Expr &)> emits = [&](const Expr &e) {
    const std::string *c = CalleeName(e);
    if (c && (*c == "PSGridEmit" || *c == "PSGridEmitUtype")) return true;
    for (const auto &k : e.kids)
      if (k && emits(*k)) return true;
    return false;
  };
  std::function<bool(const Stmt &)> any = [&](const Stmt &s) {
    if (s.expr && emits(*s.expr)) return true;
    if (s.cond && emits(*s.cond)) return true;
    if (s.init && any(*s.init)) return true;
    for (const auto &d : s.decls)
      if (d.init && emits(*d.init)) return true;
    for (const auto &k : s.kids)
      if (k && any(*k)) return true;
    return false;
  };
  return any(*f.body);
}

namespace {

bool IsAssignment(const std::string &op) {
  return op == "=" || op == "+=" || op == "-=" || op == "*=" || op == "/=" ||
         op == "%=" || op == "&=" || op == "|=" || op == "^=" ||
         op == "<<=" || op == ">>=";
}

bool NamesIndex(const KernelInfo &k, const std::string &n) {
  for (const auto &i : k.index)
    if (i == n) return true;
  return false;
}

bool RebindsIndex(const KernelInfo &k, const Expr &e) {
  const Expr *l = nullptr;
  if (e.kind == EK::kBinary && IsAssignment(e.op)) l = e.kids[0].get();
  if ((e.kind == EK::kUnary || e.kind == EK::kPostfix) &&
      (e.op == "++" || e.op == "--" || e.op == "&"))
    l = e.kids[0].get();
  while (l && l->kind == EK::kParen) l = l->kids[0].get();
  if (l && l->kind == EK::kIdent && NamesIndex(k, l->text)) return true;
  for (const auto &c : e.kids)
    if (c && RebindsIndex(k, *c)) return true;
  return false;
}

bool RebindsIndex(const KernelInfo &k, const Stmt &s) {
  for (const auto &d : s.decls) {
    if (NamesIndex(k, d.name)) return true;
    if (d.init && RebindsIndex(k, *d.init)) return true;
  }
  for (const Expr *e : {s.expr.get(), s.cond.get(), s.step.get()})
    if (e && RebindsIndex(k, *e)) return true;
  if (s.init && RebindsIndex(k, *s.init)) return true;
  for (const auto &c : s.kids)
    if (c && RebindsIndex(k, *c)) return true;
  return false;
}

}  // namespace

// An index parameter denotes the point's coordinate only while the stencil
// never redeclares, assigns, increments or takes the address of its name
// (the reference resolves every reference to its declaration,
// translator/stencil_analysis.cc:45-49); otherwise gets that use it are
// irregular.
bool Translator::RegularIndex(const KernelInfo &k, const Expr &e, int axis,
                              int *off) {
  if (k.index_rebound) return false;
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
    if (*c == "PSGridEmitUtype") {
      // PSGridEmitUtype(g.member[i]..., v) (physis_user.h): the grid is the
      // root of the member expression
      if (e.kids.size() != 3) Fail(k.name, "PSGridEmitUtype takes two arguments");
      const Expr *g = e.kids[1].get();
      while (g->kind == EK::kIndex || g->kind == EK::kMember ||
             g->kind == EK::kParen)
        g = g->kids[0].get();
      if (g->kind != EK::kIdent || k.GridIndex(g->text) < 0 ||
          !k.grids[k.GridIndex(g->text)].user)
        Fail(k.name, "PSGridEmitUtype on something that is not a member of a "
                     "grid of user-defined points");
      k.grids[k.GridIndex(g->text)].emitted = true;
    } else if (get || getp || *c == "PSGridEmit" || *c == "PSGridDim") {
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

namespace {

// Splitting a grid of user-defined points into member grids.  The runtime
// stores the points member by member (one dense plane per scalar member and
// per component of an array member), so when a stencil touches a grid g of
// points only through
//   PSGridGet(g, i...).m, PSGridGet(g, i...).a[c0][c1]  (literal components)
//   PSGridEmitUtype(g.m, v), PSGridEmitUtype(g.a[c0][c1], v)
//   PSGridGet(g, i...), PSGridEmit(g, v)  (whole points of a struct of
//                                          scalars, read and written member
//                                          by member)
//   PSGridDim(g, d)
// each plane it touches is itself a 3-D grid of the member's scalar type, and
// the stencil can run on the 2.5D kernel.  UserSplit checks a body for that
// shape and rewrites it onto the plane grids g__m, g__a_c0_c1.
struct UserSplit {
  struct Part {
    int member, comp;  // comp: row-major component of an array member, or 0
    std::string name;
  };
  const std::string g, uname;
  const std::vector<StructMember> &members;
  bool ok = true;
  std::vector<Part> used;  // first-use order

  UserSplit(const std::string &name, const std::string &point_type,
            const std::vector<StructMember> &ms)
      : g(name), uname(point_type), members(ms) {}

  static bool IsScalarType(const std::string &t) {
    return t == "float" || t == "double" || t == "int" || t == "long";
  }
  static const Expr *Unp(const Expr *e) {
    while (e && e->kind == EK::kParen) e = e->kids[0].get();
    return e;
  }
  bool IsG(const Expr *e) const {
    e = Unp(e);
    return e && e->kind == EK::kIdent && e->text == g;
  }
  static const Expr *GetOf(const Expr *e) {
    e = Unp(e);
    const std::string *c = e ? CalleeName(*e) : nullptr;
    return c && (*c == "PSGridGet" || *c == "PSGridGetPeriodic") &&
                   e->kids.size() >= 2
               ? e
               : nullptr;
  }
  // e = B.m or B.a[c0][c1]... with literal components: the member access's
  // base B (through *base) and the plane it names (*m, *comp); false if e is
  // no such form or names no scalar plane of the point type.
  bool Plane(const Expr *e, const Expr **base, int *m, int *comp) const {
    std::vector<long> idx;
    e = Unp(e);
    while (e && e->kind == EK::kIndex) {
      const Expr *i = Unp(e->kids[1].get());
      if (!i || i->kind != EK::kLiteral) return false;
      char *end = nullptr;
      const long v = std::strtol(i->text.c_str(), &end, 10);
      if (!end || *end || v < 0) return false;
      idx.insert(idx.begin(), v);
      e = Unp(e->kids[0].get());
    }
    if (!e || e->kind != EK::kMember || e->op != ".") return false;
    for (size_t i = 0; i < members.size(); ++i) {
      const StructMember &sm = members[i];
      if (sm.name != e->text || !IsScalarType(sm.type) ||
          sm.extents.size() != idx.size())
        continue;
      long c = 0;
      for (size_t d = 0; d < idx.size(); ++d) {
        if (idx[d] >= sm.extents[d]) return false;
        c = c * sm.extents[d] + idx[d];
      }
      *base = e->kids[0].get();
      *m = (int)i;
      *comp = (int)c;
      return true;
    }
    return false;
  }
  // the plane of a get of g: PSGridGet(g, i...).m / .a[c0]...
  bool GetPlane(const Expr *e, const Expr **get, int *m, int *comp) const {
    const Expr *base = nullptr;
    if (!Plane(e, &base, m, comp)) return false;
    *get = GetOf(base);
    return *get && IsG((*get)->kids[1].get());
  }
  // the plane of g an emit names: g.m / g.a[c0]...
  bool EmitPlane(const Expr *e, int *m, int *comp) const {
    const Expr *base = nullptr;
    return Plane(e, &base, m, comp) && IsG(base);
  }
  int Use(int m, int comp) {
    for (size_t i = 0; i < used.size(); ++i)
      if (used[i].member == m && used[i].comp == comp) return (int)i;
    std::string n = g + "__" + members[m].name;
    if (!members[m].extents.empty()) {
      int c = comp;
      std::vector<int> idx(members[m].extents.size());
      for (size_t d = idx.size(); d-- > 0;) {
        idx[d] = c % members[m].extents[d];
        c /= members[m].extents[d];
      }
      for (int i : idx) n += "_" + std::to_string(i);
    }
    used.push_back({m, comp, n});
    return (int)used.size() - 1;
  }
  // A whole point is read or written member by member: every member must
  // be a scalar, and the expressions that are repeated per member (the
  // indices of a get, the value of an emit) must be free of side effects.
  bool Whole(const std::vector<bool> *use = nullptr) {
    for (auto &m : members)
      if (!m.extents.empty() || !IsScalarType(m.type)) return false;
    for (size_t i = 0; i < members.size(); ++i)
      if (!use || (*use)[i]) Use((int)i, 0);
    return true;
  }
  static bool Pure(const Expr &e) {
    switch (e.kind) {
      case EK::kCall: {
        const std::string *c = CalleeName(e);
        if (!c || *c != "PSGridDim") return false;
        break;
      }
      case EK::kUnary:
      case EK::kPostfix:
        if (e.op == "++" || e.op == "--") return false;
        break;
      case EK::kBinary:
        if (!e.op.empty() && e.op.back() == '=' && e.op != "==" &&
            e.op != "!=" && e.op != "<=" && e.op != ">=")
          return false;
        break;
      case EK::kRaw:
        return false;
      default:
        break;
    }
    for (const auto &kid : e.kids)
      if (kid && !Pure(*kid)) return false;
    return true;
  }
  static ExprP Clone(const Expr &e) {
    auto c = std::make_unique<Expr>(e.kind);
    c->op = e.op;
    c->text = e.text;
    c->type = e.type;
    c->line = e.line;
    for (const auto &kid : e.kids) c->kids.push_back(kid ? Clone(*kid) : nullptr);
    return c;
  }
  static ExprP Ident(const std::string &n) {
    auto e = std::make_unique<Expr>(EK::kIdent);
    e->text = n;
    return e;
  }
  static ExprP Call(const std::string &callee) {
    auto e = std::make_unique<Expr>(EK::kCall);
    e->kids.push_back(Ident(callee));
    return e;
  }

  void Check(const Expr &e) {
    if (!ok) return;
    const Expr *get = nullptr;
    int m, comp;
    if ((e.kind == EK::kMember || e.kind == EK::kIndex) &&
        GetPlane(&e, &get, &m, &comp)) {
      // PSGridGet(g, i...).m / .a[c]...: a get of one plane
      Use(m, comp);
      for (size_t i = 2; i < get->kids.size(); ++i) Check(*get->kids[i]);
      return;
    }
    if (e.kind == EK::kIndex) {
      // an index chain over a member of a get of g that names no plane (a
      // runtime component, a scalar member indexed): not split
      const Expr *b = Unp(&e);
      while (b && b->kind == EK::kIndex) b = Unp(b->kids[0].get());
      if (b && b->kind == EK::kMember) {
        const Expr *call = GetOf(b->kids[0].get());
        if (call && IsG(call->kids[1].get())) {
          ok = false;
          return;
        }
      }
    }
    if (const std::string *c = CalleeName(e)) {
      if (*c == "PSGridEmitUtype" && e.kids.size() == 3 &&
          EmitPlane(e.kids[1].get(), &m, &comp)) {
        Use(m, comp);
        Check(*e.kids[2]);
        return;
      }
      if (*c == "PSGridDim" && e.kids.size() == 3 && IsG(e.kids[1].get())) {
        Check(*e.kids[2]);
        return;
      }
      if (*c == "PSGridEmit" && e.kids.size() == 3 && IsG(e.kids[1].get())) {
        if (!Pure(*e.kids[2]) || !Whole()) {
          ok = false;
          return;
        }
        Check(*e.kids[2]);
        return;
      }
    }
    if (const Expr *call = GetOf(&e)) {
      if (IsG(call->kids[1].get())) {  // a whole point read
        for (size_t i = 2; i < call->kids.size(); ++i)
          if (!Pure(*call->kids[i])) ok = false;
        if (!ok || !Whole()) {
          ok = false;
          return;
        }
        for (size_t i = 2; i < call->kids.size(); ++i) Check(*call->kids[i]);
        return;
      }
    }
    if (e.kind == EK::kIdent && e.text == g) {
      ok = false;  // g in any other use: a helper call, an alias, a member
      return;      // access that names no scalar plane
    }
    for (const auto &kid : e.kids)
      if (kid) Check(*kid);
  }
  // A local point initialised by a whole point read, `struct T v =
  // PSGridGet(g, i...)`, that is used only through its members reads just
  // those members' planes (the others are set to 0 and never read).
  // points: such variables and the members they use; a name declared twice
  // or used whole (passed, assigned, its address taken) uses every member.
  std::map<std::string, std::vector<bool>> points;
  bool IsWholeGet(const Declarator &d) const {
    const Expr *get = d.init ? GetOf(d.init.get()) : nullptr;
    return get && get == Unp(d.init.get()) && IsG(get->kids[1].get()) &&
           d.prefix.empty() && d.suffix.empty() && !d.ctor_init;
  }
  void ScanDecls(const Stmt &st, std::map<std::string, int> &seen) {
    for (const auto &d : st.decls) {
      ++seen[d.name];
      if (IsWholeGet(d)) points[d.name].assign(members.size(), false);
    }
    for (const auto &kid : st.kids)
      if (kid) ScanDecls(*kid, seen);
    if (st.init) ScanDecls(*st.init, seen);
  }
  void ScanUses(const Expr &e, const Expr *parent) {
    if (e.kind == EK::kIdent && points.count(e.text)) {
      std::vector<bool> &u = points[e.text];
      int m = -1;
      if (parent && parent->kind == EK::kMember && parent->op == "." &&
          parent->kids[0].get() == &e)
        for (size_t i = 0; i < members.size(); ++i)
          if (members[i].name == parent->text) m = (int)i;
      if (m >= 0)
        u[m] = true;
      else
        u.assign(members.size(), true);
    }
    for (const auto &kid : e.kids)
      if (kid) ScanUses(*kid, &e);
  }
  void ScanUses(const Stmt &st) {
    for (const auto &d : st.decls)
      if (d.init) ScanUses(*d.init, nullptr);
    for (const Expr *e : {st.expr.get(), st.cond.get(), st.step.get()})
      if (e) ScanUses(*e, nullptr);
    if (st.init) ScanUses(*st.init);
    for (const auto &kid : st.kids)
      if (kid) ScanUses(*kid);
  }
  void ScanPoints(const Stmt &body) {
    std::map<std::string, int> seen;
    ScanDecls(body, seen);
    for (auto &it : seen)
      if (it.second > 1) points.erase(it.first);
    ScanUses(body);
  }
  // the members a whole point read initialising d reads
  std::vector<bool> PointUse(const Declarator &d) const {
    auto it = points.find(d.name);
    if (!IsWholeGet(d) || it == points.end())
      return std::vector<bool>(members.size(), true);
    return it->second;
  }

  void Check(const Stmt &st) {
    for (const auto &d : st.decls) {
      if (d.name == g) ok = false;  // shadowed
      if (d.init && IsWholeGet(d)) {
        // a whole point read into a local point: the members it uses
        const Expr *get = GetOf(d.init.get());
        for (size_t i = 2; i < get->kids.size(); ++i)
          if (!Pure(*get->kids[i])) ok = false;
        const std::vector<bool> use = PointUse(d);
        if (!ok || !Whole(&use)) {
          ok = false;
          return;
        }
        for (size_t i = 2; i < get->kids.size(); ++i) Check(*get->kids[i]);
        continue;
      }
      if (d.init) Check(*d.init);
    }
    for (const Expr *e : {st.expr.get(), st.cond.get(), st.step.get()})
      if (e) Check(*e);
    if (st.init) Check(*st.init);
    for (const auto &kid : st.kids)
      if (kid) Check(*kid);
  }

  void Rewrite(ExprP &e) {
    Expr &x = *e;
    const Expr *get = nullptr;
    int m, comp;
    if ((x.kind == EK::kMember || x.kind == EK::kIndex) &&
        GetPlane(&x, &get, &m, &comp)) {
      ExprP call = Clone(*get);
      call->kids[1] = Ident(used[Use(m, comp)].name);
      for (size_t i = 2; i < call->kids.size(); ++i) Rewrite(call->kids[i]);
      e = std::move(call);
      return;
    }
    if (const std::string *c = CalleeName(x)) {
      if (*c == "PSGridEmitUtype" && x.kids.size() == 3 &&
          EmitPlane(x.kids[1].get(), &m, &comp)) {
        x.kids[0]->text = "PSGridEmit";
        x.kids[1] = Ident(used[Use(m, comp)].name);
        Rewrite(x.kids[2]);
        return;
      }
      if (*c == "PSGridDim" && x.kids.size() == 3 && IsG(x.kids[1].get())) {
        x.kids[1] = Ident(used[0].name);
        Rewrite(x.kids[2]);
        return;
      }
      if (*c == "PSGridEmit" && x.kids.size() == 3 && IsG(x.kids[1].get())) {
        // PSGridEmit(g, v) -> (PSGridEmit(g__m0, (v).m0), ...)
        Rewrite(x.kids[2]);
        ExprP chain;
        for (size_t m = 0; m < members.size(); ++m) {
          ExprP call = Call("PSGridEmit");
          call->kids.push_back(Ident(used[Use((int)m, 0)].name));
          auto val = std::make_unique<Expr>(EK::kParen);
          val->kids.push_back(Clone(*x.kids[2]));
          auto mem = std::make_unique<Expr>(EK::kMember);
          mem->op = ".";
          mem->text = members[m].name;
          mem->kids.push_back(std::move(val));
          call->kids.push_back(std::move(mem));
          if (!chain) {
            chain = std::move(call);
          } else {
            auto comma = std::make_unique<Expr>(EK::kBinary);
            comma->op = ",";
            comma->kids.push_back(std::move(chain));
            comma->kids.push_back(std::move(call));
            chain = std::move(comma);
          }
        }
        auto paren = std::make_unique<Expr>(EK::kParen);
        paren->kids.push_back(std::move(chain));
        e = std::move(paren);
        return;
      }
    }
    if (const Expr *get = GetOf(&x)) {
      if (IsG(get->kids[1].get())) {
        // a whole point read: __PSUdtMake_<type>(PSGridGet(g__m0, i...), ...)
        ExprP make = Call("__PSUdtMake_" + uname);
        for (size_t m = 0; m < members.size(); ++m) {
          ExprP one = Clone(*get);
          one->kids[1] = Ident(used[Use((int)m, 0)].name);
          for (size_t i = 2; i < one->kids.size(); ++i) Rewrite(one->kids[i]);
          make->kids.push_back(std::move(one));
        }
        e = std::move(make);
        return;
      }
    }
    for (auto &kid : x.kids)
      if (kid) Rewrite(kid);
  }
  void Rewrite(Stmt &st) {
    for (auto &d : st.decls) {
      if (d.init && IsWholeGet(d)) {
        // __PSUdtMake_<type>(PSGridGet(g__m, i...) or 0 per member)
        const std::vector<bool> use = PointUse(d);
        const Expr *get = GetOf(d.init.get());
        ExprP make = Call("__PSUdtMake_" + uname);
        for (size_t m = 0; m < members.size(); ++m) {
          if (!use[m]) {
            auto zero = std::make_unique<Expr>(EK::kLiteral);
            zero->text = "0";
            make->kids.push_back(std::move(zero));
            continue;
          }
          ExprP one = Clone(*get);
          one->kids[1] = Ident(used[Use((int)m, 0)].name);
          for (size_t i = 2; i < one->kids.size(); ++i) Rewrite(one->kids[i]);
          make->kids.push_back(std::move(one));
        }
        d.init = std::move(make);
        continue;
      }
      if (d.init) Rewrite(d.init);
    }
    for (ExprP *e : {&st.expr, &st.cond, &st.step})
      if (*e) Rewrite(*e);
    if (st.init) Rewrite(*st.init);
    for (auto &kid : st.kids)
      if (kid) Rewrite(*kid);
  }
};

}  // namespace

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
      if (IsUserGridType(t)) {
        g.user = true;
        g.uname = t.substr(8);  // PSGrid<n>D<uname>
        g.members = FindStruct(elem);
        if (!g.members)
          Fail(k.name, "grid " + g.name + ": the members of its point type " +
                           elem + " are unknown (a struct of scalars and "
                           "arrays of scalars is required)");
        g.ps_type = "PS_USER";
        g.dev_view = "__PSUdtView_" + g.uname;
      }
      k.order.push_back({'g', (int)k.grids.size()});
      k.grids.push_back(g);
    } else {
      k.order.push_back({'s', (int)k.scalars.size()});
      k.scalars.push_back({f.params[i].type, f.params[i].name});
    }
  }
  // 3-D grids of user-defined points used only member-wise become member
  // grids (the stencil body is rewritten onto them)
  if (k.ndims == 3) {
    const size_t nreal = k.grids.size();
    for (size_t gi = 0; gi < nreal; ++gi) {
      if (!k.grids[gi].user || k.grids[gi].dims != 3) continue;
      UserSplit us(k.grids[gi].name, k.grids[gi].uname, *k.grids[gi].members);
      us.ScanPoints(*f.body);
      us.Check(*f.body);
      if (!us.ok || us.used.empty()) continue;
      us.Rewrite(*const_cast<Function &>(f).body);
      k.grids[gi].split = true;
      for (const auto &part : us.used) {
        const StructMember &sm = (*k.grids[gi].members)[part.member];
        GridParam v;
        v.name = part.name;
        v.dims = 3;
        v.elem = sm.type;
        v.ps_type = PsTypeOf(sm.type);
        v.dev_view = "__PSGrid3D" + Cap(sm.type) + "_dev";
        v.member_of = (int)gi;
        v.member_index = part.member;
        v.member_comp = part.comp;
        k.grids.push_back(v);
      }
    }
  }
  k.index_rebound = RebindsIndex(k, *f.body);
  AnalyzeGets(k, *f.body);
  Decide(k);
  AnalyzeFusion(k);
}

// ------------------------------------------------- two-sweep fusion analysis
// A map pair k(a -> b), k(b -> a) over the whole grid can run as one pass
// that computes u' = k(a) on a tile plus a one-point halo and u'' = k(u')
// on the tile, when the result is exactly the two sweeps': k reads one grid
// (staged, halo <= 1) and emits the other unconditionally, and the u' values
// it computes at halo points outside the grid are never used -- either every
// neighbour get is periodic and the index parameters appear only in get /
// emit indices (so k at x = -1 equals k at x = n-1), or every neighbour get
// at offset -1 / +1 on an axis sits in the else branch of a test
// `idx == 0` / `idx == n - 1` with n = PSGridDim(grid, axis) (the clamped
// boundary of the reference diffusion benchmark, diffusion3d_physis.c:
// 40-51).
namespace {

const Expr *Unparen(const Expr *e) {
  while (e && e->kind == EK::kParen) e = e->kids[0].get();
  return e;
}

bool IsAssignOp(const std::string &op) {
  return op == "=" || op == "+=" || op == "-=" || op == "*=" || op == "/=" ||
         op == "%=" || op == "&=" || op == "|=" || op == "^=" ||
         op == "<<=" || op == ">>=";
}

struct FuseScan {
  const KernelInfo &k;
  std::map<std::string, int> assigns;   // assignments per identifier
  std::map<std::string, int> dim_axis;  // identifier = PSGridDim(g, axis)
  bool ok = true;
  std::string why;
  int emits = 0;

  explicit FuseScan(const KernelInfo &kk) : k(kk) {}

  void Fail(const std::string &w) {
    if (ok) why = w;
    ok = false;
  }
  static bool DimOf(const Expr &e, int *axis) {
    const Expr *x = Unparen(&e);
    const std::string *c = CalleeName(*x);
    int a;
    if (!c || *c != "PSGridDim" || x->kids.size() != 3 ||
        !ConstInt(*x->kids[2], &a))
      return false;
    *axis = a;
    return true;
  }
  // assignments and dimension variables
  void Assigns(const Expr &e) {
    if (e.kind == EK::kBinary && IsAssignOp(e.op)) {
      const Expr *l = Unparen(e.kids[0].get());
      if (l->kind == EK::kIdent) {
        assigns[l->text]++;
        int a;
        if (e.op == "=" && DimOf(*e.kids[1], &a)) dim_axis[l->text] = a;
      }
    }
    if ((e.kind == EK::kUnary || e.kind == EK::kPostfix) &&
        (e.op == "++" || e.op == "--")) {
      const Expr *l = Unparen(e.kids[0].get());
      if (l->kind == EK::kIdent) assigns[l->text] += 2;
    }
    for (const auto &c : e.kids)
      if (c) Assigns(*c);
  }
  void Assigns(const Stmt &s) {
    for (const auto &d : s.decls)
      if (d.init) {
        assigns[d.name]++;
        int a;
        if (DimOf(*d.init, &a)) dim_axis[d.name] = a;
        Assigns(*d.init);
      }
    for (const Expr *e : {s.expr.get(), s.cond.get(), s.step.get()})
      if (e) Assigns(*e);
    if (s.init) Assigns(*s.init);
    for (const auto &c : s.kids)
      if (c) Assigns(*c);
  }
  int Axis(const std::string &n) const {
    for (int a = 0; a < (int)k.index.size(); ++a)
      if (k.index[a] == n) return a;
    return -1;
  }
  // `idx_a == 0` -> (a, 0); `idx_a == n - 1` with n a dimension variable of
  // axis a (assigned once) or PSGridDim(g, a) - 1 -> (a, 1)
  bool Guard(const Expr &cond, std::pair<int, int> *g) const {
    const Expr *c = Unparen(&cond);
    if (c->kind != EK::kBinary || c->op != "==") return false;
    for (int side = 0; side < 2; ++side) {
      const Expr *l = Unparen(c->kids[side].get());
      const Expr *r = Unparen(c->kids[1 - side].get());
      if (l->kind != EK::kIdent) continue;
      const int a = Axis(l->text);
      if (a < 0) continue;
      int v;
      if (ConstInt(*r, &v) && v == 0) {
        *g = {a, 0};
        return true;
      }
      if (r->kind == EK::kBinary && r->op == "-" && ConstInt(*r->kids[1], &v) &&
          v == 1) {
        const Expr *n = Unparen(r->kids[0].get());
        int da;
        if (n->kind == EK::kIdent && dim_axis.count(n->text) &&
            assigns.at(n->text) == 1 && dim_axis.at(n->text) == a) {
          *g = {a, 1};
          return true;
        }
        if (DimOf(*n, &da) && da == a) {
          *g = {a, 1};
          return true;
        }
      }
    }
    return false;
  }
  // gets of the read grid must be guarded on every axis they step along
  void Gets(const Expr &e, std::vector<std::pair<int, int>> &facts,
            bool periodic) {
    if (e.kind == EK::kTernary && !periodic) {
      Gets(*e.kids[0], facts, periodic);
      std::pair<int, int> g;
      if (Guard(*e.kids[0], &g)) {
        Gets(*e.kids[1], facts, periodic);
        facts.push_back(g);
        Gets(*e.kids[2], facts, periodic);
        facts.pop_back();
        return;
      }
    }
    if (const std::string *c = CalleeName(e)) {
      if (*c == "PSGridEmit") {
        emits++;
        for (size_t i = 2; i < e.kids.size(); ++i) Gets(*e.kids[i], facts, periodic);
        return;
      }
      if (*c == "PSGridGet" || *c == "PSGridGetPeriodic") {
        if (!periodic) {
          std::array<int, 3> o = {0, 0, 0};
          for (int a = 0; a < 3; ++a)
            if (!RegularIndexOf(k, *e.kids[2 + a], a, &o[a]))
              return Fail("an irregular get");
          for (int a = 0; a < 3; ++a) {
            if (!o[a]) continue;
            bool have = false;
            for (auto &f : facts)
              have = have || (f.first == a && f.second == (o[a] < 0 ? 0 : 1));
            if (!have) return Fail("a neighbour get without a boundary guard");
          }
        }
        return;  // index arguments are get indices
      }
      if (*c == "PSGridDim") return;
      if (c->compare(0, 2, "PS") != 0) return Fail("a call of " + *c);
    }
    if (periodic && e.kind == EK::kIdent && Axis(e.text) >= 0)
      return Fail("an index parameter used outside get / emit indices");
    for (const auto &kid : e.kids)
      if (kid) Gets(*kid, facts, periodic);
  }
  void Gets(const Stmt &s, std::vector<std::pair<int, int>> &facts,
            bool periodic) {
    if (s.kind == SK::kIf) {
      Gets(*s.expr, facts, periodic);
      std::pair<int, int> g;
      const bool guard = !periodic && Guard(*s.expr, &g);
      if (!s.kids.empty() && s.kids[0]) Gets(*s.kids[0], facts, periodic);
      if (s.kids.size() > 1 && s.kids[1]) {
        if (guard) facts.push_back(g);
        Gets(*s.kids[1], facts, periodic);
        if (guard) facts.pop_back();
      }
      return;
    }
    for (const auto &d : s.decls)
      if (d.init) Gets(*d.init, facts, periodic);
    for (const Expr *e : {s.expr.get(), s.cond.get(), s.step.get()})
      if (e) Gets(*e, facts, periodic);
    if (s.init) Gets(*s.init, facts, periodic);
    for (const auto &c : s.kids)
      if (c) Gets(*c, facts, periodic);
  }
  static bool RegularIndexOf(const KernelInfo &k, const Expr &e, int axis,
                             int *off) {
    const Expr *x = Unparen(&e);
    const std::string &idx = k.index[axis];
    if (x->kind == EK::kIdent && x->text == idx) {
      *off = 0;
      return true;
    }
    if (x->kind == EK::kBinary && (x->op == "+" || x->op == "-")) {
      const Expr *a = Unparen(x->kids[0].get());
      int c;
      if (a->kind == EK::kIdent && a->text == idx && ConstInt(*x->kids[1], &c)) {
        *off = x->op == "+" ? c : -c;
        return true;
      }
    }
    return false;
  }
};

}  // namespace

namespace {
int CountEmits(const Expr &e) {
  int n = 0;
  if (const std::string *c = CalleeName(e))
    if (*c == "PSGridEmit") ++n;
  for (const auto &kid : e.kids)
    if (kid) n += CountEmits(*kid);
  return n;
}
int CountEmits(const Stmt &s) {
  int n = 0;
  for (const auto &d : s.decls)
    if (d.init) n += CountEmits(*d.init);
  for (const Expr *e : {s.expr.get(), s.cond.get(), s.step.get()})
    if (e) n += CountEmits(*e);
  if (s.init) n += CountEmits(*s.init);
  for (const auto &c : s.kids)
    if (c) n += CountEmits(*c);
  return n;
}
}  // namespace

// Two modes.  Whole grid: the pair covers the whole grid and the u' values
// the pass computes at halo points outside the grid are provably unused
// (FuseScan).  Domain: the pair covers a domain strictly inside the grid;
// the pass computes u' = k(u) at domain points and takes the emitted
// grid's own values elsewhere (a map leaves points outside its domain
// unchanged), so k needs no boundary analysis and may read further
// centre-only grids and return a value to reduce.
void Translator::AnalyzeFusion(KernelInfo &k) {
  k.fusable = false;
  k.fuse_domain = false;
  auto no = [&](const std::string &w) { k.fuse_why = w; };
  if (!cfg_.two_sweep) return no("HIP_TWO_SWEEP = false");
  if (!k.fast) return no("not a 2.5D run kernel");
  if (k.rows2d) return no("a 2-D map");
  if (k.index_rebound) return no("an index parameter is rebound");
  if (k.used_rb) {
    // one pass per red-black iteration: the map reads only the grid it
    // updates, at offsets of at most one point including the centre, and
    // emits it once, unconditionally, at the top of its body
    if (k.ret_type != "void") return no("a reducing red-black map");
    int gi = -1;
    for (int i = 0; i < (int)k.grids.size(); ++i) {
      const GridParam &g = k.grids[i];
      if (g.user) return no("a grid of user-defined points");
      if (!g.read && !g.emitted) continue;
      if (!(g.read && g.emitted) || gi >= 0 || g.irregular ||
          g.mode != GridParam::kStaged || g.periodic)
        return no("a red-black map reading other grids or periodic gets");
      gi = i;
    }
    if (gi < 0) return no("a red-black map reading no grid");
    const GridParam &g = k.grids[gi];
    bool centre = false;
    for (auto &o : g.offsets) centre = centre || (!o[0] && !o[1] && !o[2]);
    if (!centre) return no("a red-black map not reading its centre");
    if (g.hx > 1 || g.hy > 1 || g.zl < -1 || g.zr > 1)
      return no("halo wider than one point");
    if (CountEmits(*k.fn->body) != 1) return no("not exactly one emit");
    bool top = false;
    for (const auto &st : k.fn->body->kids)
      if (st && st->kind == SK::kExpr && st->expr) {
        const std::string *c = CalleeName(*Unparen(st->expr.get()));
        top = top || (c && *c == "PSGridEmit");
      }
    if (!top) return no("the emit is conditional");
    k.fuse_rb = true;
    k.fuse_in = k.fuse_out = gi;
    return;
  }
  int in = -1, out = -1, direct = 0;
  for (int i = 0; i < (int)k.grids.size(); ++i) {
    const GridParam &g = k.grids[i];
    if (g.user) return no("a grid of user-defined points");
    if (g.read && g.emitted) return no("a grid both read and emitted");
    if (g.emitted) {
      if (out >= 0) return no("more than one emitted grid");
      out = i;
    } else if (g.read && g.mode == GridParam::kStaged) {
      if (in >= 0) return no("more than one grid read at offsets");
      in = i;
    } else if (g.read) {
      if (g.mode != GridParam::kDirect || g.irregular)
        return no("grid " + g.name + " is read irregularly");
      ++direct;
    }
  }
  if (in < 0 || out < 0) return no("not one read grid and one emitted grid");
  const GridParam &g = k.grids[in];
  if (g.irregular) return no("the read grid is not staged");
  if (g.elem != "float" || k.grids[out].elem != "float")
    return no("not float grids");
  if (g.hx > 1 || g.hy > 1 || g.zl < -1 || g.zr > 1)
    return no("halo wider than one point");
  for (size_t i = 0; i < g.offsets.size(); ++i) {
    const auto &o = g.offsets[i];
    if ((o[0] || o[1] || o[2]) && g.offset_periodic[i] != g.periodic)
      return no("periodic and non-periodic neighbour gets mixed");
  }
  if (CountEmits(*k.fn->body) != 1) return no("not exactly one emit");
  bool top = false;
  for (const auto &st : k.fn->body->kids)
    if (st && st->kind == SK::kExpr && st->expr) {
      const std::string *c = CalleeName(*Unparen(st->expr.get()));
      top = top || (c && *c == "PSGridEmit");
    }
  if (!top) return no("the emit is conditional");
  k.fuse_in = in;
  k.fuse_out = out;
  std::string whole_why;
  if (direct == 0 && k.ret_type == "void") {
    FuseScan scan(k);
    scan.Assigns(*k.fn->body);
    std::vector<std::pair<int, int>> facts;
    scan.Gets(*k.fn->body, facts, g.periodic);
    if (scan.ok) {
      k.fusable = true;
      return;
    }
    whole_why = scan.why;
  }
  if (g.periodic)
    return no(whole_why.empty() ? "periodic gets over a domain" : whole_why);
  k.fusable = true;
  k.fuse_domain = true;
}

void Translator::Decide(KernelInfo &k) {
  for (int a = 1; a < k.ndims; ++a)
    k.compares_axis[a] = ComparesIdent(*k.fn->body, k.index[a]);
  k.fast = k.ndims == 3;
  if (k.ndims == 2) {
    // every grid the kernel touches 2-D and of a primitive type: the rows
    // view (grids as nx x 1 x ny, y marched as z) runs it on the 2.5D path
    bool rows = true;
    for (auto &g : k.grids)
      if ((g.read || g.emitted) && (g.dims != 2 || g.user)) rows = false;
    k.fast = k.rows2d = rows;
  }
  if (!k.fast)
    k.slow_reason = k.ndims == 2 ? "a 2-D stencil over grids that are not all "
                                   "2-D grids of a primitive type"
                                 : "not a 2-D or 3-D stencil";
  for (auto &g : k.grids)
    if (g.user && !g.split && k.fast) {
      k.fast = false;
      k.slow_reason = "grid " + g.name + " holds user-defined points";
    }
  if (k.used_rb) {
    // A colour sweep updates its grid in place.  REF visits the points of
    // a colour in loop order, so it equals a parallel sweep only when no
    // point reads another point of its own colour: every regular get of an
    // emitted grid other than the centre is at an odd offset sum, and none
    // is irregular.
    bool exact = true;
    for (auto &g : k.grids) {
      if (!g.emitted || !g.read) continue;
      if (g.irregular) exact = false;
      for (auto &o : g.offsets)
        if ((o[0] || o[1] || o[2]) && ((o[0] + o[1] + o[2]) % 2 == 0))
          exact = false;
    }
    k.rb_note = exact ? "in-place colour sweep reads only the other colour"
                      : "in-place colour sweep reads its own colour: results "
                        "depend on the order of updates";
    if (!exact) {
      std::fprintf(stderr, "physisc-hip: warning: %s: %s\n", k.name.c_str(),
                   k.rb_note.c_str());
      // the generic kernel's one thread per column keeps the REF order
      // along z at least
      k.fast = false;
      k.slow_reason = "red-black map reading its own colour";
    }
    // Otherwise the 2.5D kernel runs a colour as a masked in-place sweep:
    // every value it reads of the emitted grid is of the other colour
    // (never written in this sweep) or the point's own centre (written only
    // by its own lane, after the read), so staged planes, halo rows of
    // neighbouring tiles and neighbouring z chunks are read race-free.
  }
  for (auto &g : k.grids)
    if (k.fast && g.dims != (k.rows2d ? 2 : 3) && (g.read || g.emitted)) {
      // a 3-D map may read lower-dimensional grids (coefficients along an
      // axis: the reference's test_mixed-dim*); their gets read memory
      if (!k.rows2d && g.dims < 3 && !g.emitted && !g.user) continue;
      k.fast = false;
      k.slow_reason = g.emitted ? "grid " + g.name + " is emitted and not 3-D"
                                : "grid " + g.name + " is not 3-D";
    }
  if (!k.fast) return;
  if (k.rows2d)
    for (auto &g : k.grids) {
      for (auto &o : g.offsets) o = {o[0], 0, o[1]};
      if (g.dims == 2) g.dev_view = "__PSGrid3D" + Cap(g.elem) + "_dev";
    }
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
  if ((e->kind == EK::kIndex || e->kind == EK::kMember) &&
      RewriteUserMemberGet(k, e))
    return;
  for (auto &kid : e->kids)
    if (kid) RewriteKernelExpr(k, kid);
  const std::string *c = CalleeName(*e);
  if (!c) return;
  bool get = *c == "PSGridGet", getp = *c == "PSGridGetPeriodic";
  if ((get || getp) && k.Grid(e->kids[1]->text)->user) {
    // the whole point: every member loaded at the offset
    const GridParam &g = *k.Grid(e->kids[1]->text);
    e = Raw("__PSUdtLoad_" + g.uname + "(__ctx." + g.name + "_u, " +
            UserOffset(k, g, *e) + ")");
    return;
  }
  if (*c == "PSGridEmit" && k.Grid(e->kids[1]->text)->user) {
    const GridParam &g = *k.Grid(e->kids[1]->text);
    e = Raw("__PSUdtStore_" + g.uname + "(__ctx." + g.name + "_u, " +
            UserPointOffset(k, g) + ", " + Print(*e->kids[2]) + ")");
    return;
  }
  if (*c == "PSGridEmitUtype") {
    // g.member[i]... = v at the current point
    std::vector<ExprP *> idx;
    Expr *x = e->kids[1].get();
    while (x->kind == EK::kParen) x = x->kids[0].get();
    while (x->kind == EK::kIndex) {
      idx.insert(idx.begin(), &x->kids[1]);
      x = x->kids[0].get();
      while (x->kind == EK::kParen) x = x->kids[0].get();
    }
    if (x->kind != EK::kMember || x->op != ".")
      Fail(k.name, "PSGridEmitUtype needs grid.member as its first argument");
    const Expr *gx = x->kids[0].get();
    while (gx->kind == EK::kParen) gx = gx->kids[0].get();
    const GridParam &g = *k.Grid(gx->text);
    const StructMember *m = nullptr;
    for (auto &mm : *g.members)
      if (mm.name == x->text) m = &mm;
    if (!m) Fail(k.name, "point type of " + g.name + " has no member " + x->text);
    std::string comp = UserComponent(k, g, *m, idx);
    e = Raw("physis_hip::UdtStore<" + m->type + ">(__ctx." + g.name + "_u.m_" +
            m->name + ", __ctx." + g.name + "_u.n, " +
            std::to_string(m->Components()) + ", " + comp + ", " +
            UserPointOffset(k, g) + ", (" + m->type + ")(" +
            Print(*e->kids[2]) + "))");
    return;
  }
  if (get || getp) {
    const GridParam &g = *k.Grid(e->kids[1]->text);
    std::array<int, 3> o = {0, 0, 0};
    bool reg = k.fast && g.dims == (k.rows2d ? 2 : 3);
    for (int a = 0; reg && a < g.dims; ++a)
      reg = RegularIndex(k, *e->kids[2 + a], a, &o[a]);
    if (k.rows2d) o = {o[0], 0, o[1]};  // the rows view's (dx, 0, dy)
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
    if (k.Grid(n) && k.Grid(n)->user)
      e = Raw("__ctx." + n + "_u.dim[" + Print(*e->kids[2]) + "]");
    else
      e = Raw("__ctx." + n + "_dim[" + Print(*e->kids[2]) + "]");
    return;
  }
}

// Offset of PSGridGet(g, i...) / PSGridGetPeriodic on a user-point grid, its
// index arguments already rewritten.
std::string Translator::UserOffset(const KernelInfo &k, const GridParam &g,
                                   const Expr &call) {
  (void)k;
  const bool per = *CalleeName(call) == "PSGridGetPeriodic";
  if ((int)call.kids.size() - 2 != g.dims)
    Fail(k.name, "get of " + g.name + " with the wrong index count");
  std::string s = std::string("physis_hip::") +
                  (per ? "OffsetPeriodic(" : "OffsetLinear(") + "__ctx." +
                  g.name + "_u.dim, " + std::to_string(g.dims);
  for (int a = 0; a < g.dims; ++a) s += ", " + Print(*call.kids[2 + a]);
  return s + ")";
}

// Offset of the current point in a user-point grid (emits).
std::string Translator::UserPointOffset(const KernelInfo &k,
                                        const GridParam &g) const {
  if (g.dims > k.ndims)
    Fail(k.name, "emit into " + g.name + ", which has more dimensions than "
                 "the stencil");
  std::string s = "((long)" + k.index[0];
  if (g.dims >= 2) s += " + (long)" + k.index[1] + " * __ctx." + g.name + "_u.dim[0]";
  if (g.dims >= 3)
    s += " + (long)" + k.index[2] + " * __ctx." + g.name + "_u.dim[0] * __ctx." +
         g.name + "_u.dim[1]";
  return s + ")";
}

// Row-major component number of member m indexed by idx (rewritten first).
std::string Translator::UserComponent(const KernelInfo &k, const GridParam &g,
                                      const StructMember &m,
                                      std::vector<ExprP *> &idx) {
  if (idx.size() != m.extents.size())
    Fail(k.name, "member " + m.name + " of the points of " + g.name +
                 " must be indexed down to a scalar");
  std::string comp = "0";
  for (size_t i = 0; i < idx.size(); ++i) {
    RewriteKernelExpr(k, *idx[i]);
    comp = i == 0 ? "(" + Print(**idx[i]) + ")"
                  : "(" + comp + ") * " + std::to_string(m.extents[i]) + " + (" +
                        Print(**idx[i]) + ")";
  }
  return comp;
}

// PSGridGet(g, i...).member[j]... (or PSGridGetPeriodic) on a grid of
// user-defined points: a load of one component plane (the CUDA target's
// member arrays, cuda_runtime_builder.cc:262-345).
bool Translator::RewriteUserMemberGet(const KernelInfo &k, ExprP &e) {
  std::vector<ExprP *> idx;
  Expr *x = e.get();
  while (x->kind == EK::kIndex) {
    idx.insert(idx.begin(), &x->kids[1]);
    x = x->kids[0].get();
    while (x->kind == EK::kParen) x = x->kids[0].get();
  }
  if (x->kind != EK::kMember || x->op != ".") return false;
  Expr *call = x->kids[0].get();
  while (call->kind == EK::kParen) call = call->kids[0].get();
  const std::string *c = CalleeName(*call);
  if (!c || (*c != "PSGridGet" && *c != "PSGridGetPeriodic")) return false;
  const GridParam *g = k.Grid(call->kids[1]->text);
  if (!g || !g->user) return false;
  const StructMember *m = nullptr;
  for (auto &mm : *g->members)
    if (mm.name == x->text) m = &mm;
  if (!m) Fail(k.name, "point type of " + g->name + " has no member " + x->text);
  // a partly indexed array member is an array value: leave it to the whole
  // point load
  if (idx.size() < m->extents.size()) return false;
  for (size_t a = 2; a < call->kids.size(); ++a)
    RewriteKernelExpr(k, call->kids[a]);
  std::vector<ExprP *> own(idx.begin(), idx.begin() + m->extents.size());
  std::string comp = UserComponent(k, *g, *m, own);
  std::string off = UserOffset(k, *g, *call);
  std::string r = "physis_hip::UdtLoad<" + m->type + ">(__ctx." + g->name +
                  "_u.m_" + m->name + ", __ctx." + g->name + "_u.n, " +
                  std::to_string(m->Components()) + ", " + comp + ", " + off +
                  ")";
  if (idx.size() > m->extents.size())
    Fail(k.name, "member " + m->name + " indexed beyond its extents");
  e = Raw(r);
  return true;
}

std::string Translator::EmitUserType(const std::string &uname) {
  int nd = 0;
  std::string elem;
  for (int d = 1; d <= 3 && elem.empty(); ++d)
    if (IsGridTypeName("PSGrid" + std::to_string(d) + "D" + uname, &nd, &elem))
      break;
  const std::vector<StructMember> *ms = FindStruct(elem);
  if (!ms) return "";  // never a stencil's point type: nothing to emit
  std::ostringstream o;
  const std::string V = "__PSUdtView_" + uname;
  auto pstype = [](const std::string &t) {
    if (t == "float") return "PS_FLOAT";
    if (t == "double") return "PS_DOUBLE";
    if (t == "int") return "PS_INT";
    if (t == "long") return "PS_LONG";
    return "PS_USER";
  };
  o << "// ---------------------------------------------- point type " << uname
    << " (" << elem << ")\n"
    << "// Type info of the reference's REF target (reference_runtime_builder.cc:\n"
    << "// 959-1065): one entry per member, scalar size, array rank and extents.\n"
    << "static __PSGridTypeMemberInfo __PSTypeMembers_" << uname << "[] = {\n";
  for (auto &m : *ms) {
    o << "  {" << pstype(m.type) << ", sizeof(" << m.type << "), "
      << m.extents.size() << ", {";
    for (size_t i = 0; i < m.extents.size(); ++i)
      o << (i ? ", " : "") << m.extents[i];
    if (m.extents.empty()) o << "0";
    o << "}},\n";
  }
  o << "};\nstatic __PSGridTypeInfo __PSTypeInfo_" << uname << " = {PS_USER, "
    << "sizeof(" << elem << "), " << ms->size() << ", __PSTypeMembers_" << uname
    << "};\n\n";
  o << "// Device view: the first component plane of each member (component c\n"
    << "// of an array member, row-major, c * n elements further).\n"
    << "struct " << V << " {\n";
  for (auto &m : *ms) o << "  " << m.type << " *m_" << m.name << ";\n";
  o << "  long n;\n  int dim[3];\n};\n\n";
  o << "static inline " << V << " __PSUdtViewOf_" << uname << "(__PSGrid *g) {\n"
    << "  " << V << " v;\n";
  for (size_t i = 0; i < ms->size(); ++i)
    o << "  v.m_" << (*ms)[i].name << " = (" << (*ms)[i].type
      << " *)__PSGridUserMember(g, " << i << ");\n";
  o << "  v.n = (long)g->num_elms;\n"
    << "  for (int d = 0; d < 3; ++d) v.dim[d] = g->dim[d];\n  return v;\n}\n\n";
  auto access = [](const StructMember &m, const std::string &var) {
    std::string z;
    for (size_t i = 0; i < m.extents.size(); ++i) z += "[0]";
    return "(&" + var + "." + m.name + z + ")";
  };
  o << "__device__ __forceinline__ " << elem << " __PSUdtLoad_" << uname
    << "(const " << V << " &v, long off) {\n  " << elem << " r{};\n"
    << "  if (off < 0) return r;\n";
  for (auto &m : *ms) {
    if (m.extents.empty())
      o << "  r." << m.name << " = v.m_" << m.name << "[off];\n";
    else
      o << "  for (int c = 0; c < " << m.Components() << "; ++c) "
        << access(m, "r") << "[c] = v.m_" << m.name << "[c * v.n + off];\n";
  }
  o << "  return r;\n}\n\n";
  o << "__device__ __forceinline__ void __PSUdtStore_" << uname << "(const " << V
    << " &v, long off, const " << elem << " &r) {\n";
  for (auto &m : *ms) {
    if (m.extents.empty())
      o << "  v.m_" << m.name << "[off] = r." << m.name << ";\n";
    else
      o << "  for (int c = 0; c < " << m.Components() << "; ++c) v.m_" << m.name
        << "[c * v.n + off] = " << access(m, "r") << "[c];\n";
  }
  o << "}\n\n";
  // a point from its members (the member-split run kernels read whole points
  // of a struct of scalars this way)
  bool scalars = true;
  for (auto &m : *ms) scalars = scalars && m.extents.empty();
  if (scalars) {
    o << "__device__ __forceinline__ " << elem << " __PSUdtMake_" << uname << "(";
    for (size_t i = 0; i < ms->size(); ++i)
      o << (i ? ", " : "") << (*ms)[i].type << " m" << i;
    o << ") {\n  " << elem << " r;\n";
    for (size_t i = 0; i < ms->size(); ++i)
      o << "  r." << (*ms)[i].name << " = m" << i << ";\n";
    o << "  return r;\n}\n\n";
  }
  return o.str();
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
namespace {

// The host handle (a field of the stencil struct s) behind a kernel grid:
// the grid itself, or for a member grid its grid of user-defined points.
std::string HostGrid(const KernelInfo &k, const GridParam &g) {
  return "s." + (g.member_of >= 0 ? k.grids[g.member_of].name : g.name);
}

// The launch argument (device view) of a kernel grid.
std::string ViewArg(const KernelInfo &k, const GridParam &g) {
  if (g.member_of >= 0)
    return "physis_hip::MemberView<" + g.dev_view + ">(" + HostGrid(k, g) +
           ", " + std::to_string(g.member_index) + ", " +
           std::to_string(g.member_comp) + ")";
  if (g.user) return "__PSUdtViewOf_" + g.uname + "(s." + g.name + ")";
  if (k.rows2d && k.fast && g.dims == 2)
    return "physis_hip::RowView<" + g.dev_view + ">(s." + g.name + ")";
  return "physis_hip::DevView<" + g.dev_view + ">(s." + g.name + ")";
}

}  // namespace

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
  if (k.used_rb)
    o << "  int __rb;  // -1 all points, 0 red, 1 black, 2 red then black\n";
  o << "};\n\n";
  std::vector<std::string> params = {"__PSDomain __dom"}, init = {"__dom"};
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
  if (k.used_rb) init.push_back("-1");
  o << "static struct __PSStencil_" << n << " __PSStencilMap_" << n << "(\n    "
    << Join(params, ", ") << ") {\n  struct __PSStencil_" << n << " __s = {"
    << Join(init, ", ") << "};\n  return __s;\n}\n\n";
  if (k.used_rb) {
    std::vector<std::string> args;
    for (auto &pp : params) args.push_back(pp.substr(pp.find_last_of(" *") + 1));
    const char *names[] = {"Red", "Black", "RedBlack"};
    for (int c = 0; c < 3; ++c)
      o << "static struct __PSStencil_" << n << " __PSStencilMap" << names[c]
        << "_" << n << "(\n    " << Join(params, ", ") << ") {\n  struct __PSStencil_"
        << n << " __s = __PSStencilMap_" << n << "(" << Join(args, ", ")
        << ");\n  __s.__rb = " << c << ";\n  return __s;\n}\n\n";
  }
  if (k.ret_type != "void") {
    std::vector<std::string> args;
    for (auto &pp : params) args.push_back(pp.substr(pp.find_last_of(" *") + 1));
    o << "static struct __PSStencil_" << n << " __PSStencilMapReduce_" << n
      << "(\n    " << k.ret_type << " *__v, int __op, " << Join(params, ", ")
      << ") {\n  struct __PSStencil_" << n << " __s = __PSStencilMap_" << n << "("
      << Join(args, ", ") << ");\n  __s.__reduce_out = __v;\n"
      << "  __s.__reduce_op = __op;\n  return __s;\n}\n\n";
  }
  return o.str();
}

std::string Translator::EmitContext(const KernelInfo &k) {
  std::ostringstream o;
  o << "// Stencil analysis of " << k.name << " ("
    << (!k.fast    ? "generic run kernel: " + k.slow_reason
        : k.rows2d ? std::string("2.5D run kernel over the rows view nx x 1 x "
                                 "ny of its 2-D grids, y marched as z")
                   : std::string("2.5D run kernel"))
    << "):\n";
  for (auto &g : k.grids) {
    o << "//   " << g.name << "  ";
    if (g.split) {
      o << "split into one grid per member it uses\n";
      continue;
    }
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
  if (k.used_rb) o << "//   red-black: " << k.rb_note << "\n";
  if (k.fast)
    o << "//   two-sweep fusion of map pairs: "
      << (k.fuse_rb ? std::string("red-black: one pass per iteration (red, "
                                  "then black) over a domain inside the grid")
          : k.fusable ? std::string(k.fuse_domain
                                        ? "yes, over a domain inside the grid"
                                        : "yes")
                    : "no (" + k.fuse_why + ")") << "\n";
  o << "struct __PSCtx_" << k.name << " {\n";
  for (auto &g : k.grids) {
    if (g.split) continue;  // its member grids stand in for it
    if (g.user) {
      o << "  " << g.dev_view << " " << g.name << "_u;\n";
      continue;
    }
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
  for (auto &g : k.grids)
    if (!g.split) params.push_back(g.dev_view + " " + g.name);
  for (auto &s : k.scalars) params.push_back(s.type + " " + s.name);
  params.push_back("physis_hip::RunShape sh");
  params.push_back(RT + " *__restrict__ __red");
  if (k.used_rb) params.push_back("int __rb");
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
      o << "    const typename S_" << G << "::Addr " << G << "_ra = S_" << G
        << "::Prepare(" << G << ".dim, x0, yrow, lane);\n";
      o << "#pragma unroll\n    for (int k = 0; k < " << nslot - 1 << "; ++k) {\n"
        << "      " << G << "_st[0].Load(" << G << "_w, " << G << ".p, " << G
        << ".dim, " << G << "_ra, zb + " << g.zl << " + k);\n"
        << "      " << G << "_st[0].Store(" << G << "_lds + k * S_" << G
        << "::kPlane, wave, lane);\n    }\n";
      o << "#pragma unroll\n    for (int k = 0; k < PD; ++k)\n      " << G
        << "_st[k].Load(" << G << "_w, " << G << ".p, " << G << ".dim, " << G
        << "_ra, zb + " << g.zr + 1 << " + k);\n";
      o << "    int " << G << "_base = 0;\n";
    } else if (g.mode == GridParam::kDirect) {
      o << "    const int " << G << "_nx = " << G << ".dim[0], " << G
        << "_nxy = " << G << ".dim[0] * " << G << ".dim[1];\n";
      o << "    const physis_hip::BufWin<" << g.elem << "> " << G
        << "_w = physis_hip::BufWin<" << g.elem << ">::Make(" << G << ".p, "
        << G << "_nxy * " << G << ".dim[2], (long)zb * " << G
        << "_nxy, (long)ze * " << G << "_nxy);\n";
      o << "    " << g.elem << " " << G << "_dv[2][VX];\n";
      o << "    " << G
        << "_w.template Load<VX, physis_hip::kStreamLoad>(xs + drow * " << G
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
    if (g.split)
      continue;
    else if (g.user)
      o << "    ctx." << g.name << "_u = " << g.name << ";\n";
    else if (!k.rows2d && g.dims < 3)  // a lower-dimensional grid
    o << "    ctx." << g.name << "_p = " << g.name << ".p;\n"
      << "    for (int d = 0; d < 3; ++d) ctx." << g.name << "_dim[d] = d < "
      << g.dims << " ? " << g.name << ".dim[d] : 1;\n";
    else if (k.rows2d)  // the stencil sees the 2-D extents (nx, ny)
    o << "    ctx." << g.name << "_p = " << g.name << ".p;\n"
      << "    ctx." << g.name << "_dim[0] = " << g.name << ".dim[0];\n"
      << "    ctx." << g.name << "_dim[1] = " << g.name << ".dim[2];\n"
      << "    ctx." << g.name << "_dim[2] = 1;\n";
    else
    o << "    ctx." << g.name << "_p = " << g.name << ".p;\n"
      << "    for (int d = 0; d < 3; ++d) ctx." << g.name << "_dim[d] = "
      << g.name << ".dim[d];\n";
  // VX > 1 runs only when every grid's x extent is a multiple of 4 and
  // tiles start at multiples of TX: the compiler may use the low bits
  o << "    if constexpr (VX > 1) {\n"
       "      __builtin_assume((xs & (VX - 1)) == 0);\n";
  for (auto &g : k.grids)
    if (!g.user && (g.read || g.emitted) && g.dims == (k.rows2d ? 2 : 3))
      o << "      __builtin_assume((ctx." << g.name << "_dim[0] & 3) == 0);\n";
  o << "    }\n";
  // The z loop is unrolled over the prefetch depth.  When that is a
  // multiple of every ring's slot count, ring slots are compile-time
  // indices (each ring's oldest plane is in slot k % NSLOT at step k of the
  // body); unrolling further to get there costs more registers than it
  // saves (measured on the 19-point kernel), so other rings keep a running
  // base.
  const int U = PrefetchFor(k);
  bool const_slots = true;
  for (auto &g : k.grids)
    if (g.mode == GridParam::kStaged && U % (g.zr - g.zl + 2)) const_slots = false;
  o << "    constexpr int U = " << U << ";\n";
  for (auto &g : k.grids)
    if (g.mode == GridParam::kStaged && const_slots)
      o << "    (void)" << g.name << "_base;\n";
  o << "    for (int zc = zb; zc < ze; zc += U) {\n#pragma unroll\n"
       "      for (int k = 0; k < U; ++k) {\n"
       "        const int z = zc + k;\n"
       "        const int cur = k & 1, nxt = cur ^ 1;\n"
       "        (void)cur; (void)nxt;\n";
  for (auto &g : k.grids)
    if (g.mode == GridParam::kStaged && const_slots)
      o << "        const int " << g.name << "_base = k % " << g.zr - g.zl + 2
        << ";\n";
  for (auto &g : k.grids)
    if (g.mode == GridParam::kDirect)
      o << "        " << g.name
        << "_w.template Load<VX, physis_hip::kStreamLoad>(xs + drow * " << g.name
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
  // Every point of the row is evaluated (staged and centre values are in
  // registers, other gets are bounds-checked loads) and the domain test
  // only masks the stores and the reduction, so the loop has no branches.
  std::ostringstream b;
  b << "#pragma unroll\n          for (int i = 0; i < VX; ++i) {\n"
       "            const int x = xs + i;\n"
       "            const bool __in = x >= dom.local_min[0] && x < dom.local_max[0]"
    << (k.used_rb ? " &&\n                (__rb < 0 || ((x + yrow + z + __rb) & 1) == 0)"
                  : "")
    << ";\n            {\n";
  std::ostringstream &ob = b;
  {
  std::ostringstream &o = ob;
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
  if (k.rows2d) args = {"x", "z", "ctx"};  // the rows view marches y as z
  for (auto &s : k.scalars) args.push_back(s.name);
  if (red)
    o << "              const " << RT << " __r = " << n << "(" << Join(args, ", ")
      << ");\n              if constexpr (RED >= 0) __racc = __in ? "
         "physis_hip::ReduceOp<" << RT
      << ", RED < 0 ? 0 : RED>::apply(__racc, __r) : __racc;\n";
  else
    o << "              " << n << "(" << Join(args, ", ") << ");\n";
  for (auto &g : k.grids)
    if (g.emitted)
      o << "              " << g.name << "_out[i] = ctx." << g.name << "_v;\n"
        << "              " << g.name << "_mask |= (__in && ctx." << g.name
        << "_e) ? 1u << i : 0u;\n";
  o << "            }\n          }\n";
  }
  o << SplitInterior(k, "yrow", "z", b.str(), "          ");
  for (auto &g : k.grids)
    if (g.emitted)
      o << "          physis_hip::StoreMasked<" << g.elem << ", VX>(" << g.name
        << "_ew, xs + yrow * " << g.name << "_enx + z * " << g.name
        << "_enxy, " << g.name << "_out, " << g.name << "_mask);\n";
  o << "        }\n";
  for (auto &g : k.grids) {
    if (g.mode != GridParam::kStaged) continue;
    int nslot = g.zr - g.zl + 2;
    o << "        " << g.name << "_st[k % PD].Store(" << g.name
      << "_lds + physis_hip::RingSlot<" << nslot << ">(" << g.name << "_base, "
      << nslot - 1 << ") * S_" << g.name << "::kPlane, wave, lane);\n";
  }
  o << "        __syncthreads();\n";
  for (auto &g : k.grids) {
    if (g.mode != GridParam::kStaged) continue;
    int nslot = g.zr - g.zl + 2;
    o << "        " << g.name << "_st[k % PD].Load(" << g.name << "_w, " << g.name
      << ".p, " << g.name << ".dim, " << g.name << "_ra, z + " << g.zr + 1
      << " + PD);\n";
    if (!const_slots)
      o << "        " << g.name << "_base = physis_hip::RingSlot<" << nslot
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
  for (auto &g : k.grids)
    if (!g.split) params.push_back(g.dev_view + " " + g.name);
  for (auto &s : k.scalars) params.push_back(s.type + " " + s.name);
  params.push_back(RT + " *__restrict__ __red");
  if (k.used_rb) params.push_back("int __rb");
  o << "// Generic run kernel (one thread per (x, y) column marching z, all\n"
       "// gets read from memory): " << k.slow_reason << ".\n";
  o << "template <int RED>\n__global__ __launch_bounds__(256) void "
       "__PSStencilRunGeneric_" << n << "(\n    " << Join(params, ", ")
    << ") {\n";
  // 1-D maps: 256 points per block; otherwise 64 x 4 (x, y) columns
  if (k.ndims >= 2)
    o << "  const int x = dom.local_min[0] + blockIdx.x * 64 + threadIdx.x % 64;\n"
         "  const int y = dom.local_min[1] + blockIdx.y * 4 + threadIdx.x / 64;\n";
  else
    o << "  const int x = dom.local_min[0] + blockIdx.x * 256 + threadIdx.x;\n";
  if (red)
    o << "  " << RT << " __racc = 0;\n  if constexpr (RED >= 0) __racc = "
         "physis_hip::ReduceOp<" << RT << ", RED < 0 ? 0 : RED>::identity();\n";
  o << "  if (x < dom.local_max[0]" << (k.ndims >= 2 ? " && y < dom.local_max[1]" : "")
    << ") {\n    __PSCtx_" << n << " ctx;\n";
  for (auto &g : k.grids)
    if (g.split)
      continue;
    else if (g.user)
      o << "    ctx." << g.name << "_u = " << g.name << ";\n";
    else
    o << "    ctx." << g.name << "_p = " << g.name << ".p;\n"
      << "    for (int d = 0; d < 3; ++d) ctx." << g.name << "_dim[d] = "
      << (g.dims == 1 ? "d < 1 ? " + g.name + ".dim[d] : 1"
                      : g.dims == 2 ? "d < 2 ? " + g.name + ".dim[d] : 1"
                                    : g.name + ".dim[d]")
      << ";\n";
  if (k.ndims == 3)
    o << "    for (int z = dom.local_min[2]; z < dom.local_max[2]; ++z) {\n";
  else
    o << "    do {\n";
  if (k.used_rb) {
    // the REF colour loop (reference_runtime_builder.cc:585-599, 634-646)
    // starts x at min + ((min & 1) ^ ((rb + y + z) % 2)) in steps of 2,
    // i.e. visits the points with x + y + z + rb even
    std::string sum = "x";
    if (k.ndims >= 2) sum += " + y";
    if (k.ndims >= 3) sum += " + z";
    o << "      if (__rb >= 0 && ((" << sum << " + __rb) & 1)) continue;\n";
  }
  for (auto &g : k.grids)
    if (g.emitted && !g.user) o << "      ctx." << g.name << "_e = false;\n";
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
    if (!g.emitted || g.user) continue;  // user points are stored in place
    std::string off = "x";
    if (g.dims >= 2) off += " + (long)y * " + g.name + ".dim[0]";
    if (g.dims >= 3)
      off += " + (long)z * " + g.name + ".dim[0] * " + g.name + ".dim[1]";
    o << "      if (ctx." << g.name << "_e) " << g.name << ".p[" << off
      << "] = ctx." << g.name << "_v;\n";
  }
  o << (k.ndims == 3 ? "    }\n  }\n" : "    } while (0);\n  }\n");
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

std::string Translator::InteriorCond(const KernelInfo &k, const std::string &y,
                                     const std::string &z, int only) const {
  std::vector<std::string> c;
  // (a 2-D kernel in the rows view: its y is the kernel's z)
  const std::string *iv[3] = {nullptr, k.rows2d ? &z : &y, k.rows2d ? nullptr : &z};
  const int nd = k.rows2d ? 2 : 3;
  for (int a = 1; a < nd; ++a) {
    if (!k.compares_axis[a]) continue;
    for (int gi = 0; gi < (int)k.grids.size(); ++gi) {
      const GridParam &g = k.grids[gi];
      if (g.user || g.dims != nd || !(g.read || g.emitted)) continue;
      if (only >= 0 && gi != only) continue;
      c.push_back(*iv[a] + " >= 1 && " + *iv[a] + " <= ctx." + g.name + "_dim[" +
                  std::to_string(a) + "] - 2");
    }
  }
  return Join(c, " && ");
}

std::string Translator::InteriorAssume(const KernelInfo &k, const std::string &y,
                                       const std::string &z, int only) const {
  std::string s;
  const std::string *iv[3] = {nullptr, k.rows2d ? &z : &y, k.rows2d ? nullptr : &z};
  const int nd = k.rows2d ? 2 : 3;
  for (int a = 1; a < nd; ++a) {
    if (!k.compares_axis[a]) continue;
    for (int gi = 0; gi < (int)k.grids.size(); ++gi) {
      const GridParam &g = k.grids[gi];
      if (g.user || g.dims != nd || !(g.read || g.emitted)) continue;
      if (only >= 0 && gi != only) continue;
      const std::string d = "ctx." + g.name + "_dim[" + std::to_string(a) + "]";
      s += "__builtin_assume(" + *iv[a] + " >= 1); __builtin_assume(" + *iv[a] +
           " <= " + d + " - 2); __builtin_assume(" + *iv[a] +
           " != 0); __builtin_assume(" + *iv[a] + " != " + d + " - 1); ";
    }
  }
  return s;
}

// body: the element loop; it runs in two forms, the interior one under the
// assumptions that y and z lie off the first and last planes of every grid
// (so the compiler folds the body's boundary tests), chosen per wave step
// by a wave-uniform branch.
std::string Translator::SplitInterior(const KernelInfo &k, const std::string &y,
                                      const std::string &z,
                                      const std::string &body,
                                      const std::string &ind, int only) const {
  const std::string cond = cfg_.interior_split ? InteriorCond(k, y, z, only) : "";
  if (cond.empty()) return body;
  std::ostringstream o;
  o << ind << "auto __pts = [&](auto __interior) {\n"
    << ind << "  if constexpr (decltype(__interior)::value) {\n"
    << ind << "    " << InteriorAssume(k, y, z, only) << "\n" << ind << "  }\n"
    << body << ind << "};\n"
    << ind << "if (" << cond << ")\n"
    << ind << "  __pts(physis_hip::Flag<true>{});\n"
    << ind << "else\n" << ind << "  __pts(physis_hip::Flag<false>{});\n";
  return o.str();
}

// Two-sweep run kernel (temporal blocking) of a fusable kernel, its launcher
// and the run-time check that a map pair may use it.  Block: BY waves, wave
// r holds source row y0 - 2 + r; ring A (LDS) holds source planes with a
// two-point x halo, ring B the intermediate planes with a one-point halo,
// six slots each.  Interval s0 computes u'(s0) and u'(s0+1) on rows y0-1 ..
// y0+TY into ring B (the halo columns x0-1 and x0+TX by the lanes of wave
// 0, which otherwise only loads) and u''(s0-3), u''(s0-2) from ring B on the
// tile's rows, then stores source planes s0+3 and s0+4: one barrier per two
// planes.
std::string Translator::EmitFusedRunKernel(const KernelInfo &k) {
  std::ostringstream o;
  const std::string &n = k.name;
  const GridParam &gi = k.grids[k.fuse_in], &go = k.grids[k.fuse_out];
  const std::string T = gi.elem, V = gi.dev_view, I = gi.name, O = go.name;
  const std::string per = gi.periodic ? "true" : "false";
  std::vector<std::string> sp, sa;
  for (auto &sc : k.scalars) {
    sp.push_back(sc.type + " " + sc.name);
    sa.push_back(sc.name);
  }
  std::string fargs = "ctx";
  for (auto &a : sa) fargs += ", " + a;
  auto rows = RowsOf(gi);
  // A red-black map (k.fuse_rb): u' is the red sweep, u'' the black one, in
  // the REF order of one iteration; a point outside the map's domain or of
  // the other colour keeps its value (the centre value of its row), and
  // the pass tiles and writes the whole grid (gd).
  const bool rb = k.fuse_rb;
  int crow = -1;
  for (size_t r = 0; r < rows.size(); ++r)
    if (rows[r].dy == 0 && rows[r].dz == 0 && rows[r].xl <= 0 && rows[r].xr >= 0)
      crow = (int)r;
  if (rb && crow < 0) Fail(n, "red-black pass without the centre row");
  const std::string TD = rb ? "gd" : "dom";
  // the update test of a red-black pass at (x, y, z) for colour c
  auto upd = [&](const std::string &x, const std::string &y,
                 const std::string &z, int c) {
    return "(" + x + " >= dom.local_min[0] && " + x + " < dom.local_max[0] && " +
           y + " >= dom.local_min[1] && " + y + " < dom.local_max[1] && " + z +
           " >= dom.local_min[2] && " + z + " < dom.local_max[2] && ((" + x +
           " + " + y + " + " + z + " + " + std::to_string(c) + ") & 1) == 0)";
  };
  // ctx fill of the lane's point i from row arrays <pre>_r<r>
  auto fill = [&](const std::string &pre, const std::string &ind) {
    std::ostringstream f;
    for (size_t i = 0; i < gi.offsets.size(); ++i) {
      const auto &off = gi.offsets[i];
      for (size_t r = 0; r < rows.size(); ++r)
        if (rows[r].dy == off[1] && rows[r].dz == off[2])
          f << ind << "ctx." << I << "[" << i << "] = " << pre << "_r" << r
            << "[i + " << off[0] - rows[r].xl << "];\n";
    }
    return f.str();
  };
  // Row arrays of the offsets' rows from a ring of six slots: the planes
  // at dz = -1, 0, 1 of the centre are in slots sm, sc, sp.
  auto slot = [](int dz) { return dz < 0 ? "sm" : (dz > 0 ? "sp" : "sc"); };
  auto read_rows = [&](const std::string &pre, const std::string &lds,
                       const std::string &S, const std::string &ind) {
    std::ostringstream f;
    for (size_t r = 0; r < rows.size(); ++r) {
      const Row &w = rows[r];
      f << ind << T << " " << pre << "_r" << r << "[VX + " << w.xr - w.xl
        << "];\n";
      f << ind << S << "::template ReadRow<" << w.xl << ", " << w.xr << ">("
        << lds << " + " << slot(w.dz) << " * " << S << "::kPlane, wave + "
        << w.dy << ", lane, " << pre << "_r" << r << ");\n";
    }
    return f.str();
  };
  o << "// Two-sweep run kernel of " << n
    << " (temporal blocking): one pass over src\n"
       "// computes u' = " << n << "(src) on the tile and a one-point halo and\n"
       "// u'' = " << n << "(u') three planes behind; dst <- u'', mid <- u' when\n"
       "// mid is not NULL.  Two planes per barrier: an interval computes u' of\n"
       "// planes s0 and s0 + 1 and u'' of s0 - 3 and s0 - 2 from rings of six\n"
       "// slots (source plane p in slot (p - zb + 2) % 6, u'(p) in slot\n"
       "// (p - zb + 1) % 6), then stores source planes s0 + 3 and s0 + 4.\n";
  if (rb)
    o << "// Red-black: one pass is one iteration (red sweep u', black sweep\n"
         "// u''); dom is the map's domain, gd the whole grid the pass tiles.\n";
  o << "template <int VX, int BY>\n__global__ __launch_bounds__(64 * BY) void __PSStencilRun2_"
    << n << "(\n    __PSDomain dom, " << (rb ? "__PSDomain gd, " : "") << V
    << " src, " << V << " dst, " << T
    << " *__restrict__ mid" << (sp.empty() ? "" : ", " + Join(sp, ", "))
    << ", physis_hip::RunShape sh) {\n";
  o << "  constexpr int PD = 2, TX = 64 * VX, TY = BY - 4;\n"
    << "  typedef physis_hip::Stage<" << T << ", VX, BY, 2, 2, 6, " << per
    << "> SA;\n"
    << "  typedef physis_hip::Stage<" << T << ", VX, BY, 1, 1, 6, false> SB;\n"
    << "  __shared__ __attribute__((aligned(16))) " << T << " a_lds[SA::kLdsElems];\n"
    << "  __shared__ __attribute__((aligned(16))) " << T << " b_lds[SB::kLdsElems];\n"
    << "  const int lane = threadIdx.x % 64;\n"
       "  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);\n"
       "  const physis_hip::TileCoord tc = physis_hip::TileOf(sh.tiles_x, sh.tiles_y);\n"
       "  const int x0 = (sh.x_first_tile + tc.bx) * TX;\n"
    << "  const int y0 = " << TD << ".local_min[1] + tc.by * TY;\n"
    << "  const int zb = " << TD << ".local_min[2] + tc.bz * sh.zchunk;\n"
    << "  const int ze = min(zb + sh.zchunk, " << TD << ".local_max[2]);\n"
       "  if (zb >= ze) return;  // uniform over the block\n"
       "  const int yrow = y0 - 2 + wave;\n"
       "  const bool mid_row = wave >= 1 && wave < BY - 1;\n"
    << "  const bool out_row = wave >= 2 && wave < BY - 2 && yrow < " << TD
    << ".local_max[1];\n"
       "  const int xs = x0 + lane * VX;\n"
       "  const int nx = src.dim[0], nxy = src.dim[0] * src.dim[1];\n"
       "  const int n = nxy * src.dim[2];\n"
       "  // tiles start at multiples of TX and two-sweep runs need nx % 4 == 0\n"
       "  __builtin_assume((xs & (VX - 1)) == 0);\n"
       "  __builtin_assume((nx & (VX - 1)) == 0);\n"
    << "  const physis_hip::BufWin<" << T << "> aw = physis_hip::BufWin<" << T
    << ">::Make(src.p, n, (long)(zb - 3) * nxy, (long)(ze + 3) * nxy);\n"
    << "  const physis_hip::BufWin<" << T << "> dw = physis_hip::BufWin<" << T
    << ">::Make(dst.p, n, (long)zb * nxy, (long)ze * nxy);\n"
    << "  const physis_hip::BufWin<" << T << "> mw = physis_hip::BufWin<" << T
    << ">::Make(mid ? mid : dst.p, n, (long)zb * nxy, (long)ze * nxy);\n"
    << "  SA st[PD];\n"
       "  const typename SA::Addr ra = SA::Prepare(src.dim, x0, yrow, lane);\n"
       "#pragma unroll\n  for (int k = 0; k < 4; ++k) {\n"
       "    st[0].Load(aw, src.p, src.dim, ra, zb - 2 + k);\n"
       "    st[0].Store(a_lds + k * SA::kPlane, wave, lane);\n  }\n"
       "#pragma unroll\n  for (int k = 0; k < PD; ++k)\n"
       "    st[k].Load(aw, src.p, src.dim, ra, zb + 2 + k);\n"
       "  __syncthreads();\n"
    << "  __PSCtx_" << n << " ctx;\n"
    << "  ctx." << I << "_p = src.p;\n  ctx." << O << "_p = dst.p;\n"
    // (a fusable pair's grids have equal extents: one set of extents lets
    // the compiler fold the interior tests of both grids into one)
    << "  for (int d = 0; d < 3; ++d) {\n    ctx." << I << "_dim[d] = src.dim[d];\n"
    << "    ctx." << O << "_dim[d] = src.dim[d];\n  }\n";
  // u'(s) of the wave's row, source planes s - 1, s, s + 1 in A slots sm,
  // sc, sp, into B slot bo; the halo columns x0 - 1 and x0 + TX of u' rows
  // 1 .. BY-2 by wave 0 (which otherwise only loads), lane 2(r-1) + c
  o << "  auto __uprime = [&](const int s, const int sm, const int sc, "
       "const int sp, const int bo) {\n"
       "    if (mid_row && s <= ze) {\n";
  o << read_rows("a", "a_lds", "SA", "      ");
  o << "      " << T << " uo[VX];\n";
  {
    std::ostringstream b;
    b << "#pragma unroll\n      for (int i = 0; i < VX; ++i) {\n";
    b << fill("a", "        ");
    b << "        ctx." << O << "_e = false;\n"
      << "        " << n << "(xs + i, yrow, s, " << fargs << ");\n";
    if (rb)
      b << "        uo[i] = " << upd("(xs + i)", "yrow", "s", 0) << " ? ctx." << O
        << "_v : a_r" << crow << "[i + " << -rows[crow].xl << "];\n      }\n";
    else
      b << "        uo[i] = ctx." << O << "_v;\n      }\n";
    o << SplitInterior(k, "yrow", "s", b.str(), "      ", k.fuse_in);
  }
  o << "      SB::WriteRow(b_lds + bo * SB::kPlane, wave, lane, uo);\n";
  o << "      if (mid && out_row && s >= zb && s < ze) {\n"
       "        unsigned m = 0;\n"
       "#pragma unroll\n        for (int i = 0; i < VX; ++i)\n"
       "          if (xs + i >= dom.local_min[0] && xs + i < dom.local_max[0]) m |= 1u << i;\n"
    << "        physis_hip::StoreMasked<" << T
    << ", VX>(mw, xs + yrow * nx + s * nxy, uo, m);\n      }\n    }\n";
  o << "    if (wave == 0 && lane < 2 * (BY - 2) && s <= ze) {\n"
       "      const int hr = 1 + (lane >> 1);\n"
       "      const int px = (lane & 1) ? TX : -1;\n";
  for (size_t i = 0; i < gi.offsets.size(); ++i) {
    const auto &off = gi.offsets[i];
    o << "      ctx." << I << "[" << i << "] = a_lds[" << slot(off[2])
      << " * SA::kPlane + (hr + " << off[1]
      << ") * SA::kRow + SA::kPad + px + " << off[0] << "];\n";
  }
  o << "      ctx." << O << "_e = false;\n"
    << "      " << n << "(x0 + px, y0 - 2 + hr, s, " << fargs << ");\n";
  if (rb)
    o << "      b_lds[bo * SB::kPlane + hr * SB::kRow + SB::kPad + px] = "
      << upd("(x0 + px)", "(y0 - 2 + hr)", "s", 0) << " ? ctx." << O
      << "_v : a_lds[sc * SA::kPlane + hr * SA::kRow + SA::kPad + px];\n    }\n  };\n";
  else
    o << "      b_lds[bo * SB::kPlane + hr * SB::kRow + SB::kPad + px] = ctx."
      << O << "_v;\n    }\n  };\n";
  // u''(z) of the wave's row, u'(z - 1), u'(z), u'(z + 1) in B slots sm,
  // sc, sp
  o << "  auto __udouble = [&](const int z, const int sm, const int sc, "
       "const int sp) {\n"
       "    if (out_row && z >= zb && z < ze) {\n";
  o << read_rows("b", "b_lds", "SB", "      ");
  // every point of the row is computed (the values come from LDS); the
  // store mask keeps the grid's columns
  o << "      " << T << " uu[VX];\n      unsigned m = 0;\n";
  {
    std::ostringstream b;
    b << "#pragma unroll\n      for (int i = 0; i < VX; ++i) {\n";
    b << fill("b", "        ");
    b << "        ctx." << O << "_e = false;\n"
      << "        " << n << "(xs + i, yrow, z, " << fargs << ");\n";
    if (rb)
      b << "        uu[i] = " << upd("(xs + i)", "yrow", "z", 1) << " ? ctx." << O
        << "_v : b_r" << crow << "[i + " << -rows[crow].xl << "];\n";
    else
      b << "        uu[i] = ctx." << O << "_v;\n";
    b << "        m |= (xs + i >= " << TD << ".local_min[0] && xs + i < " << TD
      << ".local_max[0]) ? 1u << i : 0u;\n      }\n";
    o << SplitInterior(k, "yrow", "z", b.str(), "      ", k.fuse_in);
  }
  o << "      physis_hip::StoreMasked<" << T
    << ", VX>(dw, xs + yrow * nx + z * nxy, uu, m);\n    }\n  };\n";
  // three intervals per iteration, so that ring slots are compile-time
  o << "  for (int t3 = 0;; t3 += 3) {\n#pragma unroll\n"
       "    for (int j = 0; j < 3; ++j) {\n"
       "      const int s0 = zb - 1 + 2 * (t3 + j);\n"
       "      if (s0 - 3 > ze - 1) return;  // uniform over the block: every u'' done\n"
       "      __uprime(s0, (2 * j) % 6, (2 * j + 1) % 6, (2 * j + 2) % 6, (2 * j) % 6);\n"
       "      __uprime(s0 + 1, (2 * j + 1) % 6, (2 * j + 2) % 6, (2 * j + 3) % 6, "
       "(2 * j + 1) % 6);\n"
       "      __udouble(s0 - 3, (2 * j + 2) % 6, (2 * j + 3) % 6, (2 * j + 4) % 6);\n"
       "      __udouble(s0 - 2, (2 * j + 3) % 6, (2 * j + 4) % 6, (2 * j + 5) % 6);\n"
       "      st[0].Store(a_lds + ((2 * j + 4) % 6) * SA::kPlane, wave, lane);\n"
       "      st[1].Store(a_lds + ((2 * j + 5) % 6) * SA::kPlane, wave, lane);\n"
       "      __syncthreads();\n"
       "      st[0].Load(aw, src.p, src.dim, ra, s0 + 5);\n"
       "      st[1].Load(aw, src.p, src.dim, ra, s0 + 6);\n"
       "    }\n  }\n}\n\n";
  // launcher.  Launch shape: 4 points per lane in 12-wave blocks (8
  // output rows; the two six-slot rings take 152 KB of LDS, one block per
  // CU); measured on the 512^3 clamped kernel: 239 us per launch, against
  // 374 us in 8-wave blocks, and 257 us for one plane per barrier in
  // 16-wave blocks with four-slot rings.
  const int vx = 4, by = 12;
  o << "static constexpr int __PS2VX_" << n << " = " << vx << ", __PS2BY_"
    << n << " = " << by << ";\n";
  std::vector<std::string> largs = {"s.dom", "src", "dst", "mid"};
  if (rb) largs = {"s.dom", "gd", "src", "dst", "mid"};
  for (auto &a : sa) largs.push_back("s." + a);
  largs.push_back("sh");
  o << "static void __PSLaunch2_" << n << "(const struct __PSStencil_" << n
    << " &s, const " << V << " &src, const " << V << " &dst, " << T
    << " *mid, hipStream_t stream) {\n"
    << "  constexpr int VX = __PS2VX_" << n << ", BY = __PS2BY_" << n << ";\n"
    << (rb ? "  const __PSDomain gd = physis_hip::WholeDomain(s." + I + ");\n" : "")
    << "  static int per_cu = physis_hip::BlocksPerCU(\n"
       "      (const void *)__PSStencilRun2_" << n << "<VX, BY>, 64 * BY);\n"
    << "  const physis_hip::RunShape sh = physis_hip::PlanRun("
    << (rb ? "gd" : "s.dom") << ", 64 * VX, BY - 4, per_cu,\n"
       "      physis_hip::WindowChunkLimit({s." << I << "}, 3));\n"
    << "  if (sh.blocks == 0) return;\n"
    << "  hipLaunchKernelGGL((__PSStencilRun2_" << n << "<VX, BY>), "
       "dim3(sh.blocks), dim3(64 * BY), 0, stream, "
    << Join(largs, ", ") << ");\n}\n\n";
  // run-time check
  if (rb) {
    // one map of both colours over a domain at least one point inside the
    // grid (so no point of the domain reads past a face, and the halo
    // points the pass computes outside the grid are never updated)
    o << "// The red-black map s0 may run as one pass per iteration: both "
         "colours, a domain\n// strictly inside a grid whose x extent is a "
         "multiple of 4.\n";
    o << "static bool __PSFusableRB_" << n << "(const struct __PSStencil_" << n
      << " &s0) {\n"
      << "  static const bool off = getenv(\"PHYSIS_HIP_NO_FUSE\") != NULL;\n"
      << "  const __PSGrid *a = s0." << I << ";\n"
      << "  if (off || !a || s0.__rb != 2 || a->num_dims != 3 || a->dim[0] % 4) "
         "return false;\n"
      << "  for (int d = 0; d < 3; ++d)\n"
         "    if (s0.dom.local_min[d] < 1 || s0.dom.local_max[d] > a->dim[d] - 1 ||\n"
         "        s0.dom.local_min[d] >= s0.dom.local_max[d]) return false;\n"
      << "  return true;\n}\n\n";
    return o.str();
  }
  o << "// The pair (s0, s1) is " << n << "(a -> b), " << n
    << "(b -> a) over the whole grid with the\n// same scalars: it may run as "
       "two-sweep passes.\n";
  o << "static bool __PSFusable2_" << n << "(const struct __PSStencil_" << n
    << " &s0, const struct __PSStencil_" << n << " &s1) {\n"
    << "  static const bool off = getenv(\"PHYSIS_HIP_NO_FUSE\") != NULL;\n"
    << "  const __PSGrid *a = s0." << I << ", *b = s0." << O << ";\n"
    << "  if (off || !a || !b || a == b || s1." << I << " != b || s1." << O
    << " != a) return false;\n"
    << "  if (a->num_dims != 3 || b->num_dims != 3 || a->dim[0] % 4) return false;\n"
    << "  for (int d = 0; d < 3; ++d) {\n"
       "    if (a->dim[d] != b->dim[d]) return false;\n"
       "    if (s0.dom.local_min[d] != 0 || s0.dom.local_max[d] != a->dim[d]) return false;\n"
       "    if (s1.dom.local_min[d] != 0 || s1.dom.local_max[d] != a->dim[d]) return false;\n"
       "  }\n";
  for (auto &a : sa)
    o << "  if (__builtin_memcmp(&s0." << a << ", &s1." << a << ", sizeof(s0." << a
      << "))) return false;\n";
  o << "  return true;\n}\n\n";
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
  // a 2-D kernel on the 2.5D path runs over the rows view of its domain
  const std::string D = k.fast && k.rows2d ? "physis_hip::RowDomain(s.dom)" : "s.dom";
  std::vector<std::string> args = {D};
  for (auto &g : k.grids)
    if (!g.split) args.push_back(ViewArg(k, g));
  for (auto &s : k.scalars) args.push_back("s." + s.name);
  std::string A = Join(args, ", ");
  if (k.fast) {
    // LDS budget picks the block height (one wave for the rows view, whose
    // y extent is one row).
    int by = k.rows2d ? 1 : cfg_.block_waves;
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
    // buffer windows reach past the z chunk by `margin` planes per side
    int margin = 0;
    std::vector<std::string> wgrids;
    for (auto &g : k.grids) {
      if (!(g.read || g.emitted) || g.user) continue;
      if (g.member_of < 0 && g.dims != (k.rows2d ? 2 : 3)) continue;  // not windowed
      if (std::find(wgrids.begin(), wgrids.end(), HostGrid(k, g)) == wgrids.end())
        wgrids.push_back(HostGrid(k, g));
      if (g.mode == GridParam::kStaged)
        margin = std::max(margin, std::max(1 - g.zl, g.zr + 1));
    }
    o << "template <int VX>\nstatic physis_hip::RunShape __PSShape_" << n
      << "(const struct __PSStencil_" << n << " &s) {\n"
      << "  static int per_cu = physis_hip::BlocksPerCU(\n"
      << "      (const void *)__PSStencilRun_" << n << "<VX, " << by
      << ", -1>, 64 * " << by << ");\n"
      << "  return physis_hip::PlanRun(" << D << ", 64 * VX, " << by << " - 2 * "
      << k.hy << ", per_cu, physis_hip::WindowChunkLimit"
      << (k.rows2d ? "Rows" : "") << "({" << Join(wgrids, ", ") << "}, "
      << margin << "));\n}\n\n";
    // red-black maps: one launch per colour, red first
    const std::string RBA = k.used_rb ? ", __c" : "";
    o << "template <int VX>\nstatic int __PSLaunch_" << n << "(const struct __PSStencil_"
      << n << " &s, hipStream_t stream, void *red) {\n"
      << "  const physis_hip::RunShape sh = __PSShape_" << n << "<VX>(s);\n"
      << "  if (sh.blocks == 0) return 0;\n"
      << "  const dim3 grid(sh.blocks), block(64 * " << by << ");\n";
    if (k.used_rb)
      o << "  for (int __c = s.__rb == 2 ? 0 : s.__rb; "
           "__c <= (s.__rb == 2 ? 1 : s.__rb); ++__c) {\n";
    else
      o << "  {\n";
    if (red && k.used_reduce) {
      o << "  switch (red ? s.__reduce_op : -1) {\n";
      for (const char *op : {"PS_MAX", "PS_MIN", "PS_SUM", "PS_PROD"})
        o << "    case " << op << ":\n      hipLaunchKernelGGL((__PSStencilRun_" << n
          << "<VX, " << by << ", " << op << ">), grid, block, 0, stream, " << A
          << ", sh, (" << RT << " *)red" << RBA << ");\n      break;\n";
      o << "    default:\n";
    } else {
      o << "  (void)red;\n  {\n";
    }
    o << "      hipLaunchKernelGGL((__PSStencilRun_" << n << "<VX, " << by
      << ", -1>), grid, block, 0, stream, " << A << ", sh, (" << RT
      << " *)NULL" << RBA << ");\n  }\n  }\n  return sh.blocks;\n}\n\n";
    // vector width: every 3-D grid's x extent a multiple of 4
    // (the member grids of one grid of user-defined points share its checks)
    std::vector<std::string> conds, checks, emit_checks;
    auto add = [](std::vector<std::string> &v, const std::string &x) {
      if (std::find(v.begin(), v.end(), x) == v.end()) v.push_back(x);
    };
    for (auto &g : k.grids) {
      if (g.read || g.emitted) {
        // (lower-dimensional grids are read through bounds-checked gets)
        if (g.member_of >= 0 || g.dims == (k.rows2d ? 2 : 3))
          add(conds, HostGrid(k, g) + "->dim[0] % 4 == 0");
        add(checks, g.member_of >= 0
                        ? "physis_hip::CheckUserGrid3D(" + HostGrid(k, g) + ")"
                        : "physis_hip::CheckGrid3D(s." + g.name + ", " +
                              g.ps_type + ")");
      }
      if (g.emitted)
        add(emit_checks,
            "physis_hip::CheckEmitDomain(s.dom, " + HostGrid(k, g) + ")");
    }
    std::string vec = conds.empty() ? "true" : Join(conds, " && ");
    o << "static int __PSBlocks_" << n << "(const struct __PSStencil_" << n
      << " &s) {\n  return (" << vec << ") ? __PSShape_" << n
      << "<4>(s).blocks : __PSShape_" << n << "<1>(s).blocks;\n}\n\n";
    o << "static int __PSRun_" << n << "(const struct __PSStencil_" << n
      << " &s, hipStream_t stream, void *red) {\n";
    for (auto &c : checks) o << "  " << c << ";\n";
    // A periodic neighbour across an odd extent has the point's own colour
    // (x = n-1 and x = 0 share a parity): REF's sequential sweep then reads
    // a value it already updated, a colour sweep in parallel the old one.
    if (k.used_rb) {
      std::vector<std::string> odd;
      for (auto &g : k.grids) {
        if (!g.read || !g.emitted || g.dims != 3) continue;
        for (size_t i = 0; i < g.offsets.size(); ++i)
          for (int a = 0; a < 3; ++a)
            if (g.offset_periodic[i] && g.offsets[i][a])
              add(odd, HostGrid(k, g) + "->dim[" + std::to_string(a) + "] % 2");
      }
      if (!odd.empty())
        o << "  if (" << Join(odd, " || ") << ") {\n"
          << "    static bool __warned = false;\n"
          << "    if (!__warned) std::fprintf(stderr, \"[physis-hip] warning: "
          << n << ": a periodic neighbour across an odd extent has the "
             "point's own colour; colours are updated in parallel, not in "
             "the REFERENCE order\\n\");\n"
          << "    __warned = true;\n  }\n";
    }
    for (auto &c : emit_checks) o << "  " << c << ";\n";
    o << "  return (" << vec << ") ? __PSLaunch_" << n
      << "<4>(s, stream, red) : __PSLaunch_" << n << "<1>(s, stream, red);\n}\n\n";
  } else {
    o << "static dim3 __PSGridDim_" << n << "(const struct __PSStencil_" << n
      << " &s) {\n  const int nx = s.dom.local_max[0] - s.dom.local_min[0];\n"
      << "  const int ny = " << (k.ndims >= 2 ? "s.dom.local_max[1] - s.dom.local_min[1]" : "1")
      << ";\n  if (nx <= 0 || ny <= 0) return dim3(0, 0, 1);\n"
      << "  return dim3(" << (k.ndims >= 2 ? "(nx + 63) / 64, (ny + 3) / 4"
                                          : "(nx + 255) / 256, 1")
      << ", 1);\n}\n\n";
    o << "static int __PSBlocks_" << n << "(const struct __PSStencil_" << n
      << " &s) {\n  const dim3 g = __PSGridDim_" << n
      << "(s);\n  return g.x * g.y;\n}\n\n";
    o << "static int __PSRun_" << n << "(const struct __PSStencil_" << n
      << " &s, hipStream_t stream, void *red) {\n";
    for (auto &g : k.grids)
      if (g.emitted)
        o << "  physis_hip::CheckEmitDomain(s.dom, " << HostGrid(k, g)
          << ");\n";
    if (k.ndims == 3)
      o << "  if (s.dom.local_max[2] <= s.dom.local_min[2]) return 0;\n";
    o << "  const dim3 grid = __PSGridDim_" << n << "(s);\n"
      << "  if (grid.x == 0) return 0;\n";
    // red-black maps: one launch per colour, red first
    const std::string RBA = k.used_rb ? ", __c" : "";
    if (k.used_rb)
      o << "  for (int __c = s.__rb == 2 ? 0 : s.__rb; "
           "__c <= (s.__rb == 2 ? 1 : s.__rb); ++__c) {\n";
    else
      o << "  {\n";
    if (red && k.used_reduce) {
      o << "  switch (red ? s.__reduce_op : -1) {\n";
      for (const char *op : {"PS_MAX", "PS_MIN", "PS_SUM", "PS_PROD"})
        o << "    case " << op << ":\n      hipLaunchKernelGGL((__PSStencilRunGeneric_"
          << n << "<" << op << ">), grid, dim3(256), 0, stream, " << A << ", ("
          << RT << " *)red" << RBA << ");\n      break;\n";
      o << "    default:\n";
    } else {
      o << "  (void)red;\n  {\n";
    }
    o << "      hipLaunchKernelGGL((__PSStencilRunGeneric_" << n
      << "<-1>), grid, dim3(256), 0, stream, " << A << ", (" << RT
      << " *)NULL" << RBA << ");\n  }\n  }\n  return grid.x * grid.y;\n}\n\n";
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
  if (k.fusable || k.fuse_rb)
    s += k.fuse_domain ? EmitFusedDomainKernel(k) : EmitFusedRunKernel(k);
  return s;
}

// The run-function branch of a fusable pair (s0, s1) = k(a -> b), k(b -> a):
// every pair of the run is one two-sweep pass when __fuse holds.
std::string Translator::EmitFusedRun(const KernelInfo &k) {
  std::ostringstream o;
  const GridParam &gi = k.grids[k.fuse_in], &go = k.grids[k.fuse_out];
  const std::string V = gi.dev_view, T = gi.elem;
  o << "  int i0 = 0;\n  if (__fuse) {\n"
    << "    physis_hip::CheckGrid3D(s0." << gi.name << ", " << gi.ps_type << ");\n"
    << "    physis_hip::CheckGrid3D(s0." << go.name << ", " << go.ps_type << ");\n"
    << "    " << V << " va = physis_hip::DevView<" << V << ">(s0." << gi.name << ");\n"
    << "    " << V << " vb = physis_hip::DevView<" << V << ">(s0." << go.name << ");\n"
    << "    const size_t bytes = (size_t)s0." << gi.name << "->num_elms * sizeof("
    << T << ");\n"
    << "    " << V << " vt = va;\n"
    << "    vt.p = (" << T << " *)__PSHipScratch(8, bytes);\n";
  if (!k.fuse_domain) {
    // Each pass is one pair.  Even counts ping-pong a -> t -> a ...; odd
    // counts (>= 3) send the first pass to b instead, a -> b -> t -> a ...
    // (b is free until the last pass writes it).  Either way the last pass
    // reads t and writes a, and its intermediate sweep to b.
    o << "    " << V << " src = va;\n"
      << "    int j = 0;\n"
      << "    if (iter % 2) {\n"
      << "      __PSLaunch2_" << k.name << "(s0, va, vb, NULL, stream);\n"
      << "      src = vb;\n      j = 1;\n    }\n"
      << "    for (; j < iter; ++j) {\n"
      << "      const " << V << " dst = ((iter - 1 - j) % 2 == 0) ? va : vt;\n"
      << "      __PSLaunch2_" << k.name
      << "(s0, src, dst, j == iter - 1 ? vb.p : NULL, stream);\n"
      << "      src = dst;\n    }\n";
  } else {
    // Points outside the domain keep a's values in a, so the scratch grids
    // the passes write start as copies of a (stores stay inside the
    // domain).  Even counts ping-pong a -> t -> a ...; odd counts (>= 3) go
    // a -> t -> t2 -> a -> t -> a ... (b's values outside the domain are
    // read by every pass, so b cannot hold a pass).  The last pass writes a,
    // and its intermediate sweep to b.
    const bool red = k.ret_type != "void";
    o << "    " << V << " vt2 = va;\n"
      << "    physis_hip::CopyDevice(vt.p, va.p, bytes, stream);\n"
      << "    if (iter % 2) {\n"
      << "      vt2.p = (" << T << " *)__PSHipScratch(9, bytes);\n"
      << "      physis_hip::CopyDevice(vt2.p, va.p, bytes, stream);\n    }\n"
      << "    " << V << " src = va;\n"
      << "    for (int j = 0; j < iter; ++j) {\n"
      << "      " << V << " dst = (j % 2 == (iter % 2 ? 0 : 1)) ? va : vt;\n"
      << "      if (iter % 2 && j < 2) dst = j == 0 ? vt : vt2;\n"
      << "      __PSLaunch2_" << k.name << "(s0, src, dst, j == iter - 1, stream, "
      << (red ? "red0.Slot(j), red1.Slot(j)" : "NULL, NULL") << ");\n";
    if (red) o << "      red0.After(j);\n      red1.After(j);\n";
    o << "      src = dst;\n    }\n";
  }
  o << "    i0 = iter;\n  }\n";
  return o.str();
}

// Two-sweep run kernel of a map pair over a domain strictly inside the grid
// (the Himeno pair, reference examples/himeno/himenobmtxpa_physis.c:
// 379-385; the technique of examples/diffusion-benchmark/
// diffusion3d_cuda_temporal_blocking.cu:78-161).  Layout as the whole-grid
// kernel: ring A holds source planes s-1..s+1 with a two-point x halo, ring
// B the intermediate planes with a one-point halo, u'(s) is computed on
// rows y0-1 .. y0+TY (the halo columns by lanes of wave 0), u''(s-2) on the
// tile.  u' at points outside the domain is the emitted grid's own value
// (read from it), u' inside the domain is k(src); centre-only grids are
// read for u' one step ahead and again for u'' at the top of its step.
std::string Translator::EmitFusedDomainKernel(const KernelInfo &k) {
  std::ostringstream o;
  const std::string &n = k.name;
  const GridParam &gi = k.grids[k.fuse_in], &go = k.grids[k.fuse_out];
  const std::string T = gi.elem, V = gi.dev_view, I = gi.name, O = go.name;
  const bool red = k.ret_type != "void";
  const std::string RT = red ? k.ret_type : "float";
  std::vector<const GridParam *> direct;
  for (auto &g : k.grids)
    if (g.mode == GridParam::kDirect && g.read) direct.push_back(&g);
  std::vector<std::string> sp, sa;
  for (auto &sc : k.scalars) {
    sp.push_back(sc.type + " " + sc.name);
    sa.push_back(sc.name);
  }
  auto rows = RowsOf(gi);
  auto fill = [&](const std::string &pre, const std::string &ind) {
    std::ostringstream f;
    for (size_t i = 0; i < gi.offsets.size(); ++i) {
      const auto &off = gi.offsets[i];
      for (size_t r = 0; r < rows.size(); ++r)
        if (rows[r].dy == off[1] && rows[r].dz == off[2])
          f << ind << "ctx." << I << "[" << i << "] = " << pre << "_r" << r
            << "[i + " << off[0] - rows[r].xl << "];\n";
    }
    return f.str();
  };
  auto read_rows = [&](const std::string &pre, const std::string &lds,
                       const std::string &S, const std::string &base,
                       const std::string &ind) {
    std::ostringstream f;
    for (size_t r = 0; r < rows.size(); ++r) {
      const Row &w = rows[r];
      f << ind << T << " " << pre << "_r" << r << "[VX + " << w.xr - w.xl
        << "];\n";
      f << ind << S << "::template ReadRow<" << w.xl << ", " << w.xr << ">("
        << lds << " + physis_hip::RingSlot<4>(" << base << ", " << w.dz + 1
        << ") * " << S << "::kPlane, wave + " << w.dy << ", lane, " << pre
        << "_r" << r << ");\n";
    }
    return f.str();
  };
  std::vector<std::string> args = {"__PSDomain dom", V + " src", V + " dst",
                                   V + " bg", "int store_mid"};
  for (auto *g : direct) args.push_back(g->dev_view + " " + g->name);
  for (auto &x : sp) args.push_back(x);
  args.push_back("physis_hip::RunShape sh");
  args.push_back(RT + " *__restrict__ red1");
  args.push_back(RT + " *__restrict__ red2");
  std::vector<std::string> kargs_scalars;
  for (auto &a : sa) kargs_scalars.push_back(a);
  const std::string fa = kargs_scalars.empty() ? "" : ", " + Join(kargs_scalars, ", ");
  const std::string ACC = "physis_hip::ReduceOp<" + RT + ", RED < 0 ? 0 : RED>";
  o << "// Two-sweep run kernel of " << n << " over a domain strictly inside\n"
       "// the grid: one pass over src computes u' = " << n << "(src) on the\n"
       "// tile and a one-point halo (bg's own values outside the domain) and\n"
       "// u'' = " << n << "(u') two planes behind; dst <- u'', and bg <- u'\n"
       "// when store_mid.  red1 / red2: per-block partials of the two sweeps.\n";
  o << "template <int VX, int BY, int RED>\n__global__ __launch_bounds__(64 * BY) "
       "void __PSStencilRun2_" << n << "(\n    " << Join(args, ", ") << ") {\n";
  o << "  constexpr int PD = 2, TX = 64 * VX, TY = BY - 4;\n"
    << "  typedef physis_hip::Stage<" << T << ", VX, BY, 2, 2, 4, false> SA;\n"
    << "  typedef physis_hip::Stage<" << T << ", VX, BY, 1, 1, 4, false> SB;\n"
    << "  __shared__ __attribute__((aligned(16))) " << T << " a_lds[SA::kLdsElems];\n"
    << "  __shared__ __attribute__((aligned(16))) " << T << " b_lds[SB::kLdsElems];\n"
    << "  const int lane = threadIdx.x % 64;\n"
       "  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);\n"
       "  const physis_hip::TileCoord tc = physis_hip::TileOf(sh.tiles_x, sh.tiles_y);\n"
       "  const int x0 = (sh.x_first_tile + tc.bx) * TX;\n"
       "  const int y0 = dom.local_min[1] + tc.by * TY;\n"
       "  const int zb = dom.local_min[2] + tc.bz * sh.zchunk;\n"
       "  const int ze = min(zb + sh.zchunk, dom.local_max[2]);\n"
    << "  " << RT << " acc1 = 0, acc2 = 0;\n"
    << "  if constexpr (RED >= 0) acc1 = acc2 = " << ACC << "::identity();\n"
    << "  if (zb < ze) {\n"
       "    const int yrow = y0 - 2 + wave;\n"
       "    const bool mid_row = wave >= 1 && wave < BY - 1;\n"
       "    const bool out_row = wave >= 2 && wave < BY - 2 && yrow < dom.local_max[1];\n"
       "    const int xs = x0 + lane * VX;\n"
       "    const int nx = src.dim[0], nxy = src.dim[0] * src.dim[1];\n"
       "    const int n = nxy * src.dim[2];\n"
       "    if constexpr (VX > 1) {\n"
       "      __builtin_assume((xs & (VX - 1)) == 0);\n"
       "      __builtin_assume((nx & 3) == 0);\n    }\n"
       "    // u' halo columns x0 - 1 and x0 + TX of rows y0 - 1 .. y0 + TY: one\n"
       "    // point per lane of wave 0 (which otherwise only loads)\n"
       "    const bool hlane = wave == 0 && lane < 2 * (BY - 2);\n"
       "    const int hr = 1 + (lane >> 1);\n"
       "    const int px = (lane & 1) ? TX : -1;\n"
       "    const int hx = x0 + px, hy = y0 - 2 + hr;\n"
       "    const bool yin = yrow >= dom.local_min[1] && yrow < dom.local_max[1];\n"
       "    const bool hin = hy >= dom.local_min[1] && hy < dom.local_max[1] &&\n"
       "                     hx >= dom.local_min[0] && hx < dom.local_max[0];\n"
       "    bool xin_all = true;\n"
       "#pragma unroll\n    for (int i = 0; i < VX; ++i)\n"
       "      xin_all = xin_all && xs + i >= dom.local_min[0] && xs + i < dom.local_max[0];\n"
       "    const auto zin = [&](int z) { return z >= dom.local_min[2] && z < dom.local_max[2]; };\n"
    << "    const physis_hip::BufWin<" << T << "> aw = physis_hip::BufWin<" << T
    << ">::Make(src.p, n, (long)(zb - 3) * nxy, (long)(ze + 3) * nxy);\n"
    << "    const physis_hip::BufWin<" << T << "> dw = physis_hip::BufWin<" << T
    << ">::Make(dst.p, n, (long)zb * nxy, (long)ze * nxy);\n"
    << "    const physis_hip::BufWin<" << T << "> bw = physis_hip::BufWin<" << T
    << ">::Make(bg.p, n, (long)(zb - 1) * nxy, (long)(ze + 1) * nxy);\n"
    << "    const physis_hip::BufWin<" << T << "> mw = physis_hip::BufWin<" << T
    << ">::Make(bg.p, n, (long)zb * nxy, (long)ze * nxy);\n";
  for (auto *g : direct)
    o << "    const physis_hip::BufWin<" << g->elem << "> " << g->name
      << "_w = physis_hip::BufWin<" << g->elem << ">::Make(" << g->name
      << ".p, n, (long)(zb - 1) * nxy, (long)(ze + 1) * nxy);\n";
  o << "    SA st[PD];\n"
       "    const typename SA::Addr ra = SA::Prepare(src.dim, x0, yrow, lane);\n"
       "#pragma unroll\n    for (int k = 0; k < 3; ++k) {\n"
       "      st[0].Load(aw, src.p, src.dim, ra, zb - 2 + k);\n"
       "      st[0].Store(a_lds + k * SA::kPlane, wave, lane);\n    }\n"
       "#pragma unroll\n    for (int k = 0; k < PD; ++k)\n"
       "      st[k].Load(aw, src.p, src.dim, ra, zb + 1 + k);\n";
  // centre values of the direct grids: plane s is loaded one step ahead of
  // u'(s) and kept for u''(s) two steps later (u'' rows are u' rows), in
  // a ring of 4 planes indexed by the unrolled step
  for (auto *g : direct)
    o << "    " << g->elem << " " << g->name << "_c1[4][VX];\n";
  o << "    " << T << " bg_c1[4][VX];\n";
  o << "    // loads of plane z for u': the lane's VX points on mid rows, its halo\n"
       "    // point in wave 0; emitted-grid values only where a point lies\n"
       "    // outside the domain\n"
       "    auto load1 = [&](int z, int slot) {\n"
       "      if (wave == 0) {\n"
       "        const int o = hlane ? hx + hy * nx + z * nxy : physis_hip::kDrop;\n"
       "        const int ob = hlane && !(hin && zin(z)) ? o : physis_hip::kDrop;\n";
  for (auto *g : direct)
    o << "        " << g->name << "_w.template Load<1>(o, *reinterpret_cast<"
      << g->elem << " (*)[1]>(&" << g->name << "_c1[slot][0]));\n";
  o << "        bw.template Load<1>(ob, *reinterpret_cast<" << T
    << " (*)[1]>(&bg_c1[slot][0]));\n"
       "      } else {\n"
       "        const int o = mid_row ? xs + yrow * nx + z * nxy : physis_hip::kDrop;\n"
       "        const int ob = mid_row && !(yin && xin_all && zin(z)) ? o : physis_hip::kDrop;\n";
  for (auto *g : direct)
    o << "        " << g->name << "_w.template Load<VX>(o, " << g->name
      << "_c1[slot]);\n";
  o << "        bw.template Load<VX>(ob, bg_c1[slot]);\n      }\n    };\n"
       "    load1(zb - 1, 0);\n"
       "    __syncthreads();\n";
  o << "    __PSCtx_" << n << " ctx;\n";
  for (auto &g : k.grids) {
    std::string src_view = &g == &gi ? "src" : (&g == &go ? "dst" : g.name);
    o << "    ctx." << g.name << "_p = " << src_view << ".p;\n"
      << "    for (int d = 0; d < 3; ++d) ctx." << g.name << "_dim[d] = "
      << src_view << ".dim[d];\n";
  }
  o << "    for (int sc = zb - 1; sc <= ze + 1; sc += 4) {\n#pragma unroll\n"
       "      for (int k = 0; k < 4; ++k) {\n"
       "        const int s = sc + k;\n"
       "        const int abase = k, bbase = k;\n"
       "        const int cur = k, nxt = (k + 1) & 3;\n";
  if (!direct.empty()) o << "        const int old = (k + 2) & 3;\n";
  o << "        if (s > ze + 1) break;  // uniform over the block\n"
       "        load1(s + 1, nxt);\n";
  // u' on the lane's points
  o << "        if (mid_row && s <= ze) {\n";
  o << read_rows("a", "a_lds", "SA", "abase", "          ");
  o << "          " << T << " uo[VX];\n          unsigned m = 0;\n"
       "          const bool own = out_row && s >= zb && s < ze;\n"
       "#pragma unroll\n          for (int i = 0; i < VX; ++i) {\n";
  o << fill("a", "            ");
  for (auto *g : direct)
    o << "            ctx." << g->name << "_c = " << g->name << "_c1[cur][i];\n";
  o << "            ctx." << O << "_e = false;\n"
    << "            const bool in = yin && xs + i >= dom.local_min[0] && "
       "xs + i < dom.local_max[0] && zin(s);\n";
  if (red)
    o << "            const " << RT << " r = " << n << "(xs + i, yrow, s, ctx" << fa
      << ");\n            if constexpr (RED >= 0) acc1 = in && own ? " << ACC
      << "::apply(acc1, r) : acc1;\n";
  else
    o << "            " << n << "(xs + i, yrow, s, ctx" << fa << ");\n";
  o << "            uo[i] = in ? ctx." << O << "_v : bg_c1[cur][i];\n"
       "            m |= in ? 1u << i : 0u;\n          }\n";
  o << "          " << T << " *brow = b_lds + physis_hip::RingSlot<4>(bbase, 3) * "
       "SB::kPlane + wave * SB::kRow + SB::kPad + lane * VX;\n"
       "#pragma unroll\n          for (int i = 0; i < VX; ++i) brow[i] = uo[i];\n"
       "          if (store_mid && own)\n"
       "            physis_hip::StoreMasked<" << T << ", VX>(mw, "
       "xs + yrow * nx + s * nxy, uo, m);\n"
       "        }\n";
  // u' halo columns
  o << "        if (hlane && s <= ze) {\n";
  for (size_t i = 0; i < gi.offsets.size(); ++i) {
    const auto &off = gi.offsets[i];
    o << "          ctx." << I << "[" << i << "] = a_lds[physis_hip::RingSlot<4>(abase, "
      << off[2] + 1 << ") * SA::kPlane + (hr + " << off[1]
      << ") * SA::kRow + SA::kPad + px + " << off[0] << "];\n";
  }
  for (auto *g : direct)
    o << "          ctx." << g->name << "_c = " << g->name << "_c1[cur][0];\n";
  o << "          ctx." << O << "_e = false;\n"
    << "          " << (red ? "(void)" : "") << n << "(hx, hy, s, ctx" << fa << ");\n"
    << "          b_lds[physis_hip::RingSlot<4>(bbase, 3) * SB::kPlane + hr * SB::kRow + "
       "SB::kPad + px] = hin && zin(s) ? ctx." << O << "_v : bg_c1[cur][0];\n"
       "        }\n";
  // u'' on the tile
  o << "        if (out_row && s - 2 >= zb) {\n";
  o << read_rows("b", "b_lds", "SB", "bbase", "          ");
  o << "          " << T << " uu[VX];\n          unsigned m = 0;\n"
       "#pragma unroll\n          for (int i = 0; i < VX; ++i) {\n";
  o << fill("b", "            ");
  for (auto *g : direct)
    o << "            ctx." << g->name << "_c = " << g->name << "_c1[old][i];\n";
  o << "            ctx." << O << "_e = false;\n"
    << "            const bool in = xs + i >= dom.local_min[0] && xs + i < dom.local_max[0];\n";
  if (red)
    o << "            const " << RT << " r = " << n << "(xs + i, yrow, s - 2, ctx" << fa
      << ");\n            if constexpr (RED >= 0) acc2 = in ? " << ACC
      << "::apply(acc2, r) : acc2;\n";
  else
    o << "            " << n << "(xs + i, yrow, s - 2, ctx" << fa << ");\n";
  o << "            uu[i] = ctx." << O << "_v;\n"
       "            m |= in ? 1u << i : 0u;\n          }\n"
       "          physis_hip::StoreMasked<" << T
    << ", VX>(dw, xs + yrow * nx + (s - 2) * nxy, uu, m);\n        }\n"
       "        st[k % PD].Store(a_lds + physis_hip::RingSlot<4>(abase, 3) * SA::kPlane, "
       "wave, lane);\n"
       "        __syncthreads();\n"
       "        st[k % PD].Load(aw, src.p, src.dim, ra, s + 2 + PD);\n"
       "      }\n    }\n  }\n";
  o << "  if constexpr (RED >= 0) {\n"
    << "    __shared__ " << RT << " lds_r1[BY], lds_r2[BY];\n"
    << "    const " << RT << " r1 = physis_hip::BlockReduce<" << RT
    << ", RED < 0 ? 0 : RED, 64 * BY>(acc1, lds_r1);\n"
    << "    const " << RT << " r2 = physis_hip::BlockReduce<" << RT
    << ", RED < 0 ? 0 : RED, 64 * BY>(acc2, lds_r2);\n"
       "    if (threadIdx.x == 0) {\n"
       "      if (red1) red1[blockIdx.x] = r1;\n"
       "      if (red2) red2[blockIdx.x] = r2;\n    }\n"
       "  } else {\n    (void)acc1; (void)acc2; (void)red1; (void)red2;\n  }\n}\n\n";
  // launcher, block count and run-time check.  Every centre-only grid
  // holds 4 * VX values per lane (its ring of 4 planes).  Without such
  // grids the kernel is as light as the whole-grid one (16-wave blocks);
  // with many (the 19-point Himeno kernel's 12), one point per lane keeps
  // the rings within the register file at four waves per SIMD (measured at
  // 512x256x256 with the per-sweep residual: 132 Gpts/s in 12-wave blocks,
  // 125 in 16-wave blocks, 101 for two points per lane in 8-wave blocks).
  const int kVX = direct.size() > 4 ? 1 : 4;
  const int kBY = direct.empty() ? 16 : (direct.size() > 4 ? 12 : 8);
  o << "static constexpr int __PS2VX_" << n << " = " << kVX << ", __PS2BY_" << n
    << " = " << kBY << ";\n";
  std::vector<std::string> wgrids = {"s." + I, "s." + O};
  for (auto *g : direct) wgrids.push_back("s." + g->name);
  o << "template <int VX, int BY>\nstatic physis_hip::RunShape __PSShape2_" << n
    << "(const struct __PSStencil_" << n << " &s) {\n"
    << "  static int per_cu = physis_hip::BlocksPerCU(\n"
    << "      (const void *)__PSStencilRun2_" << n << "<VX, BY, -1>, 64 * BY);\n"
    << "  return physis_hip::PlanRun(s.dom, 64 * VX, BY - 4, per_cu,\n"
    << "      physis_hip::WindowChunkLimit({" << Join(wgrids, ", ") << "}, 3));\n}\n\n";
  o << "static int __PSBlocks2_" << n << "(const struct __PSStencil_" << n
    << " &s) {\n  return __PSShape2_" << n << "<__PS2VX_" << n << ", __PS2BY_"
    << n << ">(s).blocks;\n}\n\n";
  std::vector<std::string> largs = {"s.dom", "src", "dst",
                                    "physis_hip::DevView<" + V + ">(s." + O + ")",
                                    "store_mid"};
  for (auto *g : direct)
    largs.push_back("physis_hip::DevView<" + g->dev_view + ">(s." + g->name + ")");
  for (auto &a : sa) largs.push_back("s." + a);
  largs.push_back("sh");
  largs.push_back("(" + RT + " *)red1");
  largs.push_back("(" + RT + " *)red2");
  const std::string LA = Join(largs, ", ");
  o << "static void __PSLaunch2_" << n << "(const struct __PSStencil_" << n
    << " &s, const " << V << " &src, const " << V
    << " &dst, int store_mid, hipStream_t stream, void *red1, void *red2) {\n"
    << "  constexpr int VX = __PS2VX_" << n << ", BY = __PS2BY_" << n << ";\n"
    << "  const physis_hip::RunShape sh = __PSShape2_" << n << "<VX, BY>(s);\n"
    << "  if (sh.blocks == 0) return;\n"
    << "  const dim3 grid(sh.blocks), block(64 * BY);\n";
  if (red && k.used_reduce) {
    o << "  switch (red1 || red2 ? s.__reduce_op : -1) {\n";
    for (const char *op : {"PS_MAX", "PS_MIN", "PS_SUM", "PS_PROD"})
      o << "    case " << op << ":\n      hipLaunchKernelGGL((__PSStencilRun2_" << n
        << "<VX, BY, " << op << ">), grid, block, 0, stream, " << LA
        << ");\n      break;\n";
    o << "    default:\n";
  } else {
    o << "  {\n";
  }
  o << "      hipLaunchKernelGGL((__PSStencilRun2_" << n
    << "<VX, BY, -1>), grid, block, 0, stream, " << LA << ");\n  }\n}\n\n";
  o << "// The pair (s0, s1) is " << n << "(a -> b), " << n
    << "(b -> a) over one domain strictly\n// inside the grids, with the same "
       "centre-only grids (neither a nor b) and\n// scalars: it may run as "
       "two-sweep passes.\n";
  o << "static bool __PSFusable2_" << n << "(const struct __PSStencil_" << n
    << " &s0, const struct __PSStencil_" << n << " &s1) {\n"
    << "  static const bool off = getenv(\"PHYSIS_HIP_NO_FUSE\") != NULL;\n"
    << "  const __PSGrid *a = s0." << I << ", *b = s0." << O << ";\n"
    << "  if (off || !a || !b || a == b || s1." << I << " != b || s1." << O
    << " != a) return false;\n"
    << "  if (a->num_dims != 3 || b->num_dims != 3 || a->dim[0] % 4) return false;\n"
    << "  for (int d = 0; d < 3; ++d) {\n"
       "    if (a->dim[d] != b->dim[d]) return false;\n"
       "    if (s0.dom.local_min[d] != s1.dom.local_min[d] ||\n"
       "        s0.dom.local_max[d] != s1.dom.local_max[d]) return false;\n"
       "    if (s0.dom.local_min[d] < 1 || s0.dom.local_max[d] > a->dim[d] - 1 ||\n"
       "        s0.dom.local_min[d] >= s0.dom.local_max[d]) return false;\n"
       "  }\n";
  for (auto *g : direct) {
    const std::string G = g->name;
    o << "  if (s0." << G << " != s1." << G << " || !s0." << G << " || s0." << G
      << " == a || s0." << G << " == b || s0." << G << "->num_dims != 3) return false;\n"
      << "  for (int d = 0; d < 3; ++d)\n    if (s0." << G
      << "->dim[d] != a->dim[d]) return false;\n";
  }
  for (auto &a : sa)
    o << "  if (__builtin_memcmp(&s0." << a << ", &s1." << a << ", sizeof(s0." << a
      << "))) return false;\n";
  if (red)
    o << "  if ((s0.__reduce_out == NULL) != (s1.__reduce_out == NULL) ||\n"
         "      (s0.__reduce_out && s0.__reduce_op != s1.__reduce_op)) return false;\n";
  o << "  return true;\n}\n\n";
  return o.str();
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
  bool any_red = false;
  for (size_t i = 0; i < r.kernels.size(); ++i)
    any_red = any_red || kernels_.at(r.kernels[i]).ret_type != "void";
  if (any_red) o << "  physis_hip::ReduceSlots __slots;\n";
  // a fusable pair decides up front whether it runs as two-sweep passes:
  // the partials of its reducing maps then come from the fused launches
  const bool fused_pair = r.kernels.size() == 2 &&
                          r.kernels[0] == r.kernels[1] &&
                          kernels_.at(r.kernels[0]).fusable;
  if (fused_pair)
    o << "  const bool __fuse = iter >= 2 && __PSFusable2_" << r.kernels[0]
      << "(s0, s1);\n";
  for (size_t i = 0; i < r.kernels.size(); ++i) {
    const KernelInfo &k = kernels_.at(r.kernels[i]);
    std::string si = "s" + std::to_string(i);
    if (k.ret_type == "void") continue;
    std::string blocks = "__PSBlocks_" + k.name + "(" + si + ")";
    if (fused_pair && k.fuse_domain)
      blocks = "(__fuse ? __PSBlocks2_" + k.name + "(" + si + ") : " + blocks +
               ")";
    o << "  physis_hip::MapReducer red" << i << "(" << si << ".__reduce_out, "
      << si << ".__reduce_op, " << PsTypeOf(k.ret_type) << ",\n      " << si
      << ".__reduce_out ? " << blocks << " : 0, iter, __slots);\n";
  }
  const bool fuse = r.kernels.size() == 2 && r.kernels[0] == r.kernels[1] &&
                    kernels_.at(r.kernels[0]).fusable;
  if (fuse) o << EmitFusedRun(kernels_.at(r.kernels[0]));
  // a red-black map of both colours: one pass per iteration, ping-ponging
  // between the grid and a scratch copy (every pass writes the whole grid);
  // an odd count ends with a copy back
  const bool fuse_rb = r.kernels.size() == 1 && kernels_.at(r.kernels[0]).fuse_rb;
  if (fuse_rb) {
    const KernelInfo &k = kernels_.at(r.kernels[0]);
    const GridParam &g = k.grids[k.fuse_in];
    const std::string V = g.dev_view, T = g.elem;
    o << "  int i0 = 0;\n"
      << "  if (iter >= 1 && __PSFusableRB_" << k.name << "(s0)) {\n"
      << "    physis_hip::CheckGrid3D(s0." << g.name << ", " << g.ps_type << ");\n"
      << "    const " << V << " va = physis_hip::DevView<" << V << ">(s0." << g.name
      << ");\n"
      << "    const size_t bytes = (size_t)s0." << g.name << "->num_elms * sizeof("
      << T << ");\n"
      << "    " << V << " vt = va;\n"
      << "    vt.p = (" << T << " *)__PSHipScratch(8, bytes);\n"
      << "    " << V << " src = va;\n"
      << "    for (int j = 0; j < iter; ++j) {\n"
      << "      const " << V << " dst = (j % 2 == 0) ? vt : va;\n"
      << "      __PSLaunch2_" << k.name << "(s0, src, dst, NULL, stream);\n"
      << "      src = dst;\n    }\n"
      << "    if (iter % 2) physis_hip::CopyDevice(va.p, vt.p, bytes, stream);\n"
      << "    i0 = iter;\n  }\n";
  }
  o << "  for (int i = " << (fuse || fuse_rb ? "i0" : "0") << "; i < iter; ++i) {\n";
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
  if (IsMapName(c)) {
    while (call.kids[1]->kind == EK::kParen || call.kids[1]->kind == EK::kCast)
      call.kids[1] = std::move(call.kids[1]->kids[0]);
    // __PSStencilMap_k, __PSStencilMapRed_k, ...
    callee->text = "__" + c + "_" + call.kids[1]->text;
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
      if (ac && (IsMapName(*ac) || *ac == "PSStencilMapReduce")) {
        const Expr *kn = (IsMapName(*ac) ? a.kids[1] : a.kids[3]).get();
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
      IsUserGridType(c.substr(0, c.size() - 3))) {
    IsGridTypeName(c.substr(0, c.size() - 3), &nd, &elem);
    if (!FindStruct(elem))
      Fail(f.name, c + ": the members of " + elem + " are unknown (a struct "
                   "of scalars and arrays of scalars is required)");
    std::string s = "physis_hip::NewUserGrid(&__PSTypeInfo_" + c.substr(8, c.size() - 11) +
                    ", " + std::to_string(nd);
    for (int a = 0; a < nd; ++a) s += ", " + Print(*e->kids[1 + a]);
    e = Raw(s + ")");
  } else if (c.size() > 3 && c.compare(c.size() - 3, 3, "New") == 0 &&
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
  } else if (c == "PSGridGet" || c == "PSGridGetPeriodic") {
    // a get in host code (reference translator.cc:170-190): one element
    // read back; periodic gets are rejected there too
    if (c == "PSGridGetPeriodic")
      Fail(f.name, "PSGridGetPeriodic is not allowed in host code");
    std::string el = GridElemOf(*e->kids[1], f);
    if (el.empty() || el.compare(0, 7, "struct ") == 0 || !IsGridTypeName(
        "PSGrid3D" + Cap(el), nullptr, nullptr))
      Fail(f.name, "PSGridGet in host code on a grid of unknown or "
                   "user-defined type");
    std::string s = "physis_hip::GetElem<" + el + ">(" + Print(*e->kids[1]);
    for (size_t a = 2; a < e->kids.size(); ++a) s += ", " + Print(*e->kids[a]);
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
    if (kernels_.count(f.name) || device_helpers_.count(f.name)) continue;
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
  std::set<std::string> emitted_kernels, user_types;
  for (auto &it : prog_.items) {
    if (it.kind == TopItem::kInclude) continue;
    if (it.kind == TopItem::kTokens) {
      if (it.cxx)
        close_c();
      else
        open_c();
      out << PrintTokens(it.tokens) << "\n\n";
      if (!it.user_grid.empty() && !user_types.count(it.user_grid)) {
        user_types.insert(it.user_grid);
        std::string u = EmitUserType(it.user_grid);
        if (!u.empty()) {
          close_c();
          out << u;
        }
      }
      continue;
    }
    Function &f = *it.func;
    if (kernels_.count(f.name)) {
      close_c();
      out << EmitKernel(kernels_[f.name]);
      emitted_kernels.insert(f.name);
      continue;
    }
    if (IsUnusedStencil(f)) {
      out << "// stencil function " << f.name
          << ": not used by any map, not emitted\n\n";
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
