// Recursive-descent parser for the Physis C subset, and the printer that
// turns (rewritten) trees back into C/C++ text.
#include "parser.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>

namespace physisc {

namespace {

const std::set<std::string> &BuiltinTypeWords() {
  static const std::set<std::string> s = {
      "void", "char", "short", "int", "long", "float", "double", "signed",
      "unsigned", "_Bool", "bool", "const", "volatile", "static", "extern",
      "register", "inline", "auto", "struct", "union", "enum", "restrict",
      "__restrict__", "__restrict"};
  return s;
}

// Typedef names that come from headers the translator does not read.
const std::set<std::string> &LibraryTypeNames() {
  static const std::set<std::string> s = {
      "size_t", "ssize_t", "ptrdiff_t", "FILE", "int8_t", "int16_t",
      "int32_t", "int64_t", "uint8_t", "uint16_t", "uint32_t", "uint64_t",
      "intptr_t", "uintptr_t", "PSIndex", "PSVectorInt", "PSPoint",
      "PSDomain1D", "PSDomain2D", "PSDomain3D", "__PSDomain", "PSStencil",
      "PSType", "__PSStopwatch", "va_list"};
  return s;
}

}  // namespace

namespace {

std::map<std::string, std::pair<int, std::string>> &UserGridTypes() {
  static std::map<std::string, std::pair<int, std::string>> m;
  return m;
}

std::map<std::string, std::vector<StructMember>> &Structs() {
  static std::map<std::string, std::vector<StructMember>> m;
  return m;
}

}  // namespace

void RegisterUserGridType(const std::string &name, int dims,
                          const std::string &elem) {
  UserGridTypes()["PSGrid" + std::to_string(dims) + "D" + name] = {dims, elem};
}

bool IsUserGridType(const std::string &s) {
  return UserGridTypes().count(s) > 0;
}

const std::vector<StructMember> *FindStruct(const std::string &type) {
  auto it = Structs().find(type);
  return it == Structs().end() ? nullptr : &it->second;
}

bool IsGridTypeName(const std::string &s, int *dims, std::string *elem) {
  auto u = UserGridTypes().find(s);
  if (u != UserGridTypes().end()) {
    if (dims) *dims = u->second.first;
    if (elem) *elem = u->second.second;
    return true;
  }
  // PSGrid{1,2,3}D{Float,Double,Int,Long}
  if (s.size() < 10 || s.compare(0, 6, "PSGrid") != 0) return false;
  char d = s[6];
  if (d < '1' || d > '3' || s[7] != 'D') return false;
  std::string e = s.substr(8);
  std::string c;
  if (e == "Float") c = "float";
  else if (e == "Double") c = "double";
  else if (e == "Int") c = "int";
  else if (e == "Long") c = "long";
  else return false;
  if (dims) *dims = d - '0';
  if (elem) *elem = c;
  return true;
}

Parser::Parser(std::vector<Token> toks, std::string file)
    : t_(std::move(toks)), file_(std::move(file)) {}

[[noreturn]] void Parser::Fail(const std::string &msg) const {
  std::fprintf(stderr, "%s:%d: physisc-hip: %s (near '%s')\n", file_.c_str(),
               Peek().line, msg.c_str(), Peek().text.c_str());
  std::exit(1);
}

const Token &Parser::Peek(int k) const {
  size_t i = pos_ + k;
  return i < t_.size() ? t_[i] : t_.back();
}

bool Parser::Accept(const char *s) {
  if (Peek().Is(s) && Peek().kind != Tok::kString && Peek().kind != Tok::kChar) {
    ++pos_;
    return true;
  }
  return false;
}

void Parser::Expect(const char *s) {
  if (!Accept(s)) Fail(std::string("expected '") + s + "'");
}

bool Parser::IsTypeName(const std::string &s) const {
  return BuiltinTypeWords().count(s) || LibraryTypeNames().count(s) ||
         typedefs_.count(s) || IsGridTypeName(s, nullptr, nullptr);
}

bool Parser::StartsDeclaration() const {
  const Token &a = Peek();
  if (a.kind != Tok::kIdent) return false;
  if (BuiltinTypeWords().count(a.text)) return true;
  if (!IsTypeName(a.text)) {
    // "T x" with an unknown T is still a declaration: two identifiers in a
    // row cannot start an expression.
    const Token &b = Peek(1);
    return b.kind == Tok::kIdent && !IsKeyword(b.text);
  }
  const Token &b = Peek(1);
  return b.kind == Tok::kIdent || b.Is("*") || b.Is("(") || b.Is("&");
}

bool Parser::IsKeyword(const std::string &s) {
  static const std::set<std::string> k = {
      "if", "else", "for", "while", "do", "return", "break", "continue",
      "switch", "case", "default", "goto", "sizeof", "typedef"};
  return k.count(s) > 0;
}

// Base type: specifiers and qualifiers up to the first declarator.
std::string Parser::ParseBaseType() {
  std::string ty;
  bool have_core = false;
  while (true) {
    const Token &a = Peek();
    if (a.kind != Tok::kIdent) break;
    if (a.text == "struct" || a.text == "union" || a.text == "enum") {
      ty += (ty.empty() ? "" : " ") + a.text;
      ++pos_;
      if (Peek().kind == Tok::kIdent) {
        ty += " " + Peek().text;
        ++pos_;
      }
      if (Peek().Is("{")) {  // inline body: keep it as text
        if (ty.compare(0, 6, "struct") == 0) {
          ty += " " + ParseStructBody(ty);
        } else {
          int depth = 0;
          do {
            if (Peek().Is("{")) ++depth;
            if (Peek().Is("}")) --depth;
            ty += " " + Peek().text;
            ++pos_;
          } while (depth > 0 && Peek().kind != Tok::kEnd);
        }
      }
      have_core = true;
      continue;
    }
    if (BuiltinTypeWords().count(a.text)) {
      ty += (ty.empty() ? "" : " ") + a.text;
      ++pos_;
      if (a.text != "const" && a.text != "volatile" && a.text != "static" &&
          a.text != "extern" && a.text != "register" && a.text != "inline" &&
          a.text != "restrict")
        have_core = true;
      continue;
    }
    if (!have_core && (IsTypeName(a.text) ||
                       (Peek(1).kind == Tok::kIdent && !IsKeyword(a.text)) ||
                       (Peek(1).Is("*") && IsTypeName(a.text)))) {
      ty += (ty.empty() ? "" : " ") + a.text;
      ++pos_;
      have_core = true;
      continue;
    }
    break;
  }
  if (ty.empty()) Fail("expected a type");
  return ty;
}

std::string Parser::ParseStructBody(const std::string &tag) {
  // Members are declarations "type name[e1][e2], ...;" of scalar types with
  // constant extents; anything else leaves the struct unregistered (it can
  // still be used, but not as the point type of a grid).
  std::string text;
  size_t start = pos_;
  Expect("{");
  std::vector<StructMember> members;
  bool plain = true;
  while (!Peek().Is("}") && Peek().kind != Tok::kEnd) {
    size_t p0 = pos_;
    std::string base;
    while (Peek().kind == Tok::kIdent && BuiltinTypeWords().count(Peek().text)) {
      if (Peek().text != "const" && Peek().text != "volatile")
        base += (base.empty() ? "" : " ") + Peek().text;
      ++pos_;
    }
    if (base.empty() || base == "struct" || base == "union" || base == "enum") {
      plain = false;
      pos_ = p0;
      int depth = 0;  // skip to the end of this member declaration
      while (Peek().kind != Tok::kEnd && !(depth == 0 && Peek().Is(";"))) {
        if (Peek().Is("{")) ++depth;
        if (Peek().Is("}")) --depth;
        ++pos_;
      }
      Accept(";");
      continue;
    }
    do {
      StructMember m;
      m.type = base;
      if (Peek().Is("*")) plain = false;
      while (Peek().Is("*")) ++pos_;
      if (Peek().kind != Tok::kIdent) Fail("expected a member name");
      m.name = Peek().text;
      ++pos_;
      while (Accept("[")) {
        const Token &n = Peek();
        if (n.kind != Tok::kNumber) {
          plain = false;
          while (!Peek().Is("]") && Peek().kind != Tok::kEnd) ++pos_;
        } else {
          m.extents.push_back((int)std::strtol(n.text.c_str(), nullptr, 0));
          ++pos_;
        }
        Expect("]");
      }
      members.push_back(m);
    } while (Accept(","));
    Expect(";");
  }
  Expect("}");
  for (size_t i = start; i < pos_; ++i) text += (i > start ? " " : "") + t_[i].text;
  if (plain && !members.empty() && tag != "struct") Structs()[tag] = members;
  return text;
}

Declarator Parser::ParseDeclarator(bool allow_init) {
  Declarator d;
  while (Peek().Is("*") || Peek().Is("&") || Peek().Is("const") ||
         Peek().Is("restrict") || Peek().Is("__restrict__")) {
    d.prefix += (Peek().text == "*" || Peek().text == "&")
                    ? Peek().text
                    : " " + Peek().text + " ";
    ++pos_;
  }
  if (Peek().kind != Tok::kIdent) Fail("expected a name");
  d.name = Peek().text;
  ++pos_;
  while (Peek().Is("[")) {
    int depth = 0;
    do {
      if (Peek().Is("[")) ++depth;
      if (Peek().Is("]")) --depth;
      d.suffix += Peek().text;
      ++pos_;
    } while (depth > 0);
  }
  if (allow_init && Accept("=")) {
    if (Peek().Is("{"))
      d.init = ParseInitList();
    else
      d.init = ParseAssign();
  } else if (allow_init && d.suffix.empty() && Peek().Is("(") &&
             !Peek(1).Is(")")) {
    // C++ direct initialisation "T v(a, b)"
    ++pos_;
    d.init = ParseExpr();
    d.ctor_init = true;
    Expect(")");
  }
  return d;
}

ExprP Parser::ParseInitList() {
  Expect("{");
  auto e = std::make_unique<Expr>(EK::kInitList);
  while (!Peek().Is("}")) {
    if (Peek().Is("{"))
      e->kids.push_back(ParseInitList());
    else
      e->kids.push_back(ParseAssign());
    if (!Accept(",")) break;
  }
  Expect("}");
  return e;
}

// ---------------------------------------------------------- expressions
ExprP Parser::ParseExpr() {
  ExprP l = ParseAssign();
  while (Peek().Is(",")) {
    ++pos_;
    auto b = std::make_unique<Expr>(EK::kBinary);
    b->op = ",";
    b->kids.push_back(std::move(l));
    b->kids.push_back(ParseAssign());
    l = std::move(b);
  }
  return l;
}

ExprP Parser::ParseAssign() {
  ExprP l = ParseTernary();
  static const std::set<std::string> ops = {"=",  "+=", "-=", "*=", "/=", "%=",
                                            "&=", "^=", "|=", "<<=", ">>="};
  if (Peek().kind == Tok::kPunct && ops.count(Peek().text)) {
    std::string op = Peek().text;
    ++pos_;
    auto b = std::make_unique<Expr>(EK::kBinary);
    b->op = op;
    b->kids.push_back(std::move(l));
    b->kids.push_back(ParseAssign());
    return b;
  }
  return l;
}

ExprP Parser::ParseTernary() {
  ExprP c = ParseBinary(0);
  if (Accept("?")) {
    auto t = std::make_unique<Expr>(EK::kTernary);
    t->kids.push_back(std::move(c));
    t->kids.push_back(ParseExpr());
    Expect(":");
    t->kids.push_back(ParseTernary());
    return t;
  }
  return c;
}

int Parser::BinaryPrec(const std::string &op) {
  static const char *levels[][6] = {
      {"||"}, {"&&"}, {"|"}, {"^"}, {"&"}, {"==", "!="},
      {"<", ">", "<=", ">="}, {"<<", ">>"}, {"+", "-"}, {"*", "/", "%"}};
  for (int i = 0; i < 10; ++i)
    for (const char *s : levels[i])
      if (s && op == s) return i + 1;
  return 0;
}

ExprP Parser::ParseBinary(int min_prec) {
  ExprP l = ParseUnary();
  while (true) {
    const Token &t = Peek();
    if (t.kind != Tok::kPunct) break;
    int p = BinaryPrec(t.text);
    if (p == 0 || p <= min_prec) break;
    std::string op = t.text;
    ++pos_;
    ExprP r = ParseBinary(p);
    auto b = std::make_unique<Expr>(EK::kBinary);
    b->op = op;
    b->kids.push_back(std::move(l));
    b->kids.push_back(std::move(r));
    l = std::move(b);
  }
  return l;
}

bool Parser::LooksLikeCast() const {
  if (!Peek().Is("(")) return false;
  const Token &a = Peek(1);
  if (a.kind != Tok::kIdent || !IsTypeName(a.text)) return false;
  // (T) or (T *...) followed by an operand
  size_t k = 2;
  if ((a.text == "struct" || a.text == "union" || a.text == "enum") &&
      Peek(2).kind == Tok::kIdent)
    k = 3;  // the tag
  while (true) {
    const Token &b = Peek(k);
    if (b.Is(")")) return true;
    if (b.Is("*") || (b.kind == Tok::kIdent && IsTypeName(b.text))) {
      ++k;
      continue;
    }
    return false;
  }
}

std::string Parser::ParseTypeName() {
  std::string ty = ParseBaseType();
  while (Peek().Is("*") || Peek().Is("const")) {
    ty += Peek().text == "*" ? "*" : " const";
    ++pos_;
  }
  return ty;
}

ExprP Parser::ParseUnary() {
  const Token &t = Peek();
  int line = t.line;
  if (t.kind == Tok::kPunct &&
      (t.text == "-" || t.text == "+" || t.text == "!" || t.text == "~" ||
       t.text == "*" || t.text == "&" || t.text == "++" || t.text == "--")) {
    std::string op = t.text;
    ++pos_;
    auto u = std::make_unique<Expr>(EK::kUnary);
    u->op = op;
    u->line = line;
    u->kids.push_back(ParseUnary());
    return u;
  }
  if (t.Is("sizeof")) {
    ++pos_;
    if (LooksLikeCast()) {
      Expect("(");
      auto s = std::make_unique<Expr>(EK::kSizeofType);
      s->type = ParseTypeName();
      Expect(")");
      return s;
    }
    auto u = std::make_unique<Expr>(EK::kUnary);
    u->op = "sizeof ";
    u->kids.push_back(ParseUnary());
    return u;
  }
  if (LooksLikeCast()) {
    Expect("(");
    auto c = std::make_unique<Expr>(EK::kCast);
    c->type = ParseTypeName();
    Expect(")");
    c->kids.push_back(ParseUnary());
    return c;
  }
  return ParsePostfix();
}

ExprP Parser::ParsePostfix() {
  ExprP e = ParsePrimary();
  while (true) {
    if (Peek().Is("(")) {
      ++pos_;
      auto c = std::make_unique<Expr>(EK::kCall);
      c->line = e->line;
      c->kids.push_back(std::move(e));
      if (!Peek().Is(")")) {
        do {
          c->kids.push_back(ParseAssign());
        } while (Accept(","));
      }
      Expect(")");
      e = std::move(c);
    } else if (Peek().Is("[")) {
      ++pos_;
      auto x = std::make_unique<Expr>(EK::kIndex);
      x->kids.push_back(std::move(e));
      x->kids.push_back(ParseExpr());
      Expect("]");
      e = std::move(x);
    } else if (Peek().Is(".") || Peek().Is("->")) {
      auto m = std::make_unique<Expr>(EK::kMember);
      m->op = Peek().text;
      ++pos_;
      if (Peek().kind != Tok::kIdent) Fail("expected a member name");
      m->text = Peek().text;
      ++pos_;
      m->kids.push_back(std::move(e));
      e = std::move(m);
    } else if (Peek().Is("++") || Peek().Is("--")) {
      auto p = std::make_unique<Expr>(EK::kPostfix);
      p->op = Peek().text;
      ++pos_;
      p->kids.push_back(std::move(e));
      e = std::move(p);
    } else {
      break;
    }
  }
  return e;
}

ExprP Parser::ParsePrimary() {
  const Token &t = Peek();
  if (t.kind == Tok::kIdent && !IsKeyword(t.text)) {
    auto e = std::make_unique<Expr>(EK::kIdent);
    e->text = t.text;
    e->line = t.line;
    ++pos_;
    return e;
  }
  if (t.kind == Tok::kNumber || t.kind == Tok::kChar) {
    auto e = std::make_unique<Expr>(EK::kLiteral);
    e->text = t.text;
    e->line = t.line;
    ++pos_;
    return e;
  }
  if (t.kind == Tok::kString) {
    auto e = std::make_unique<Expr>(EK::kLiteral);
    e->text = t.text;
    ++pos_;
    while (Peek().kind == Tok::kString) {  // adjacent literals
      e->text += " " + Peek().text;
      ++pos_;
    }
    return e;
  }
  if (t.Is("(")) {
    ++pos_;
    auto p = std::make_unique<Expr>(EK::kParen);
    p->kids.push_back(ParseExpr());
    Expect(")");
    return p;
  }
  Fail("expected an expression");
}

// ----------------------------------------------------------- statements
StmtP Parser::ParseStmt() {
  const Token &t = Peek();
  int line = t.line;
  if (t.Is("{")) return ParseCompound();
  if (t.Is(";")) {
    ++pos_;
    return std::make_unique<Stmt>(SK::kEmpty);
  }
  if (t.Is("if")) {
    ++pos_;
    auto s = std::make_unique<Stmt>(SK::kIf);
    s->line = line;
    Expect("(");
    s->expr = ParseExpr();
    Expect(")");
    s->kids.push_back(ParseStmt());
    if (Accept("else")) s->kids.push_back(ParseStmt());
    return s;
  }
  if (t.Is("for")) {
    ++pos_;
    auto s = std::make_unique<Stmt>(SK::kFor);
    s->line = line;
    Expect("(");
    if (Peek().Is(";")) {
      ++pos_;
    } else if (StartsDeclaration()) {
      s->init = ParseDecl();
    } else {
      auto e = std::make_unique<Stmt>(SK::kExpr);
      e->expr = ParseExpr();
      Expect(";");
      s->init = std::move(e);
    }
    if (!Peek().Is(";")) s->cond = ParseExpr();
    Expect(";");
    if (!Peek().Is(")")) s->step = ParseExpr();
    Expect(")");
    s->kids.push_back(ParseStmt());
    return s;
  }
  if (t.Is("while")) {
    ++pos_;
    auto s = std::make_unique<Stmt>(SK::kWhile);
    Expect("(");
    s->expr = ParseExpr();
    Expect(")");
    s->kids.push_back(ParseStmt());
    return s;
  }
  if (t.Is("do")) {
    ++pos_;
    auto s = std::make_unique<Stmt>(SK::kDo);
    s->kids.push_back(ParseStmt());
    Expect("while");
    Expect("(");
    s->expr = ParseExpr();
    Expect(")");
    Expect(";");
    return s;
  }
  if (t.Is("switch")) {
    ++pos_;
    auto s = std::make_unique<Stmt>(SK::kSwitch);
    Expect("(");
    s->expr = ParseExpr();
    Expect(")");
    s->kids.push_back(ParseStmt());
    return s;
  }
  if (t.Is("case")) {
    ++pos_;
    auto s = std::make_unique<Stmt>(SK::kCase);
    s->expr = ParseTernary();
    Expect(":");
    return s;
  }
  if (t.Is("default")) {
    ++pos_;
    Expect(":");
    return std::make_unique<Stmt>(SK::kDefault);
  }
  if (t.Is("return")) {
    ++pos_;
    auto s = std::make_unique<Stmt>(SK::kReturn);
    s->line = line;
    if (!Peek().Is(";")) s->expr = ParseExpr();
    Expect(";");
    return s;
  }
  if (t.Is("break") || t.Is("continue")) {
    ++pos_;
    Expect(";");
    return std::make_unique<Stmt>(t.text == "break" ? SK::kBreak
                                                    : SK::kContinue);
  }
  if (t.Is("goto")) {
    ++pos_;
    auto s = std::make_unique<Stmt>(SK::kGoto);
    s->label = Peek().text;
    ++pos_;
    Expect(";");
    return s;
  }
  if (t.kind == Tok::kIdent && Peek(1).Is(":") && !IsTypeName(t.text)) {
    auto s = std::make_unique<Stmt>(SK::kLabel);
    s->label = t.text;
    pos_ += 2;
    return s;
  }
  if (StartsDeclaration()) return ParseDecl();
  auto s = std::make_unique<Stmt>(SK::kExpr);
  s->line = line;
  s->expr = ParseExpr();
  Expect(";");
  return s;
}

StmtP Parser::ParseDecl() {
  auto s = std::make_unique<Stmt>(SK::kDecl);
  s->line = Peek().line;
  s->type = ParseBaseType();
  do {
    s->decls.push_back(ParseDeclarator(true));
  } while (Accept(","));
  Expect(";");
  return s;
}

StmtP Parser::ParseCompound() {
  Expect("{");
  auto s = std::make_unique<Stmt>(SK::kCompound);
  while (!Peek().Is("}")) {
    if (Peek().kind == Tok::kEnd) Fail("unterminated block");
    s->kids.push_back(ParseStmt());
  }
  Expect("}");
  return s;
}

// ----------------------------------------------------------- top level
Program Parser::ParseProgram() {
  Program prog;
  prog.file = file_;
  while (Peek().kind != Tok::kEnd) {
    size_t start = pos_;
    const Token &t = Peek();
    if (t.kind == Tok::kIdent && t.text.compare(0, 18, "__physisc_include_") == 0) {
      TopItem it;
      it.kind = TopItem::kInclude;
      it.include_line = t.text;
      ++pos_;
      Accept(";");
      prog.items.push_back(std::move(it));
      continue;
    }
    if (t.Is(";")) {
      ++pos_;
      continue;
    }
    if (t.kind == Tok::kIdent && t.text.size() == 13 &&
        t.text.compare(0, 11, "DeclareGrid") == 0 && t.text[12] == 'D' &&
        t.text[11] >= '1' && t.text[11] <= '3' && Peek(1).Is("(")) {
      // DeclareGrid<n>D(Name, type): declares the grid type PSGrid<n>DName
      // of points of `type` (physis_user.h:27-44)
      const int nd = t.text[11] - '0';
      pos_ += 2;
      if (Peek().kind != Tok::kIdent) Fail("expected a grid type name");
      std::string name = Peek().text;
      ++pos_;
      Expect(",");
      std::string elem;
      while (!Peek().Is(")") && Peek().kind != Tok::kEnd) {
        elem += (elem.empty() ? "" : " ") + Peek().text;
        ++pos_;
      }
      Expect(")");
      Accept(";");
      RegisterUserGridType(name, nd, elem);
      TopItem it;
      it.kind = TopItem::kTokens;
      it.tokens.assign(t_.begin() + start, t_.begin() + pos_);
      it.user_grid = name;
      prog.items.push_back(std::move(it));
      continue;
    }
    if (t.Is("class") || t.Is("using") || t.Is("namespace") ||
        t.Is("template")) {
      // C++ declarations outside the Physis subset are kept as written:
      // up to the ';' that ends them at brace depth 0, or to the closing
      // brace of a namespace body.  A class name becomes a type name.
      // A function template ends with its body (no ';' follows).
      const bool ns = t.Is("namespace"), tmpl = t.Is("template");
      if (t.Is("class") && Peek(1).kind == Tok::kIdent)
        typedefs_.insert(Peek(1).text);
      int depth = 0;
      while (true) {
        if (Peek().kind == Tok::kEnd) Fail("unterminated declaration");
        if (Peek().Is("{")) ++depth;
        if (Peek().Is("}") && --depth == 0 &&
            (ns || (tmpl && !Peek(1).Is(";")))) {
          ++pos_;
          break;
        }
        if (depth == 0 && Peek().Is(";")) {
          ++pos_;
          break;
        }
        ++pos_;
      }
      TopItem it;
      it.kind = TopItem::kTokens;
      it.cxx = true;
      it.tokens.assign(t_.begin() + start, t_.begin() + pos_);
      prog.items.push_back(std::move(it));
      continue;
    }
    if (t.Is("typedef")) {
      // typedef ... name ;  (register the last identifier at depth 0)
      int depth = 0;
      std::string last;
      while (!(depth == 0 && Peek().Is(";"))) {
        if (Peek().kind == Tok::kEnd) Fail("unterminated typedef");
        if (Peek().Is("{") || Peek().Is("(") || Peek().Is("[")) ++depth;
        if (Peek().Is("}") || Peek().Is(")") || Peek().Is("]")) --depth;
        if (depth == 0 && Peek().kind == Tok::kIdent) last = Peek().text;
        ++pos_;
      }
      ++pos_;
      typedefs_.insert(last);
      // typedef struct [Tag] { ... } Name;  -> Name has Tag's members
      if (Peek(0).kind != Tok::kEnd) {
        size_t q = start + 1;
        if (t_[q].Is("struct")) {
          ++q;
          std::string tag;
          if (t_[q].kind == Tok::kIdent) tag = "struct " + t_[q++].text;
          if (t_[q].Is("{")) {
            size_t save = pos_;
            pos_ = q;
            ParseStructBody(tag.empty() ? "struct __anon" : tag);
            pos_ = save;
            if (const auto *m = FindStruct(tag.empty() ? "struct __anon" : tag))
              Structs()[last] = *m;
          } else if (!tag.empty() && FindStruct(tag)) {
            Structs()[last] = *FindStruct(tag);
          }
        }
      }
      TopItem it;
      it.kind = TopItem::kTokens;
      it.tokens.assign(t_.begin() + start, t_.begin() + pos_);
      prog.items.push_back(std::move(it));
      continue;
    }
    // Declaration or function: base type, then a declarator.
    std::string base = ParseBaseType();
    if (Peek().Is(";")) {  // struct/enum definition alone
      ++pos_;
      TopItem it;
      it.kind = TopItem::kTokens;
      it.tokens.assign(t_.begin() + start, t_.begin() + pos_);
      prog.items.push_back(std::move(it));
      continue;
    }
    size_t decl_start = pos_;
    std::string stars;
    while (Peek().Is("*")) {
      stars += "*";
      ++pos_;
    }
    if (Peek().kind == Tok::kIdent && Peek(1).Is("(")) {
      auto f = std::make_unique<Function>();
      f->line = Peek().line;
      f->name = Peek().text;
      pos_ += 2;
      // storage class words stay out of the return type
      std::string ret;
      for (size_t p = 0; p < base.size();) {
        size_t q = base.find(' ', p);
        std::string w = base.substr(p, q == std::string::npos ? q : q - p);
        if (w == "static") f->is_static = true;
        else if (w == "inline") f->is_inline = true;
        else if (w != "extern") ret += (ret.empty() ? "" : " ") + w;
        if (q == std::string::npos) break;
        p = q + 1;
      }
      f->ret_type = ret + stars;
      if (Peek().Is("void") && Peek(1).Is(")")) ++pos_;
      while (!Peek().Is(")")) {
        if (Accept("...")) {
          f->variadic = true;
          break;
        }
        Param prm;
        prm.type = ParseBaseType();
        // an unnamed parameter of a prototype: "int", "char *", "int []"
        size_t q = pos_;
        while (t_[q].Is("*") || t_[q].Is("const")) ++q;
        if (t_[q].Is(",") || t_[q].Is(")") || t_[q].Is("[")) {
          for (; pos_ < q; ++pos_)
            prm.type += t_[pos_].Is("*") ? "*" : " const";
          if (Accept("[")) {
            while (!Peek().Is("]") && Peek().kind != Tok::kEnd) ++pos_;
            Expect("]");
            prm.type += "*";
          }
          f->params.push_back(prm);
          if (!Accept(",")) break;
          continue;
        }
        Declarator d = ParseDeclarator(false);
        prm.type += d.prefix.empty() ? "" : " " + d.prefix;
        prm.name = d.name;
        if (!d.suffix.empty()) prm.type += "*";  // array parameter
        f->params.push_back(prm);
        if (!Accept(",")) break;
      }
      Expect(")");
      if (Peek().Is("{")) {
        f->body = ParseCompound();
        TopItem it;
        it.kind = TopItem::kFunction;
        it.func = std::move(f);
        prog.items.push_back(std::move(it));
      } else {
        Expect(";");
        TopItem it;
        it.kind = TopItem::kTokens;
        it.tokens.assign(t_.begin() + start, t_.begin() + pos_);
        prog.items.push_back(std::move(it));
      }
      continue;
    }
    pos_ = decl_start;
    // variable declaration(s)
    do {
      Declarator d = ParseDeclarator(true);
      globals_.push_back({base + " " + d.prefix, d.name});
    } while (Accept(","));
    Expect(";");
    TopItem it;
    it.kind = TopItem::kTokens;
    it.tokens.assign(t_.begin() + start, t_.begin() + pos_);
    prog.items.push_back(std::move(it));
  }
  return prog;
}

// ------------------------------------------------------------- printer
std::string Print(const Expr &e) {
  switch (e.kind) {
    case EK::kIdent:
    case EK::kLiteral:
    case EK::kRaw:
      return e.text;
    case EK::kCall: {
      std::string s = Print(*e.kids[0]) + "(";
      for (size_t i = 1; i < e.kids.size(); ++i)
        s += (i > 1 ? ", " : "") + Print(*e.kids[i]);
      return s + ")";
    }
    case EK::kIndex:
      return Print(*e.kids[0]) + "[" + Print(*e.kids[1]) + "]";
    case EK::kMember:
      return Print(*e.kids[0]) + e.op + e.text;
    case EK::kUnary:
      return e.op + Print(*e.kids[0]);
    case EK::kPostfix:
      return Print(*e.kids[0]) + e.op;
    case EK::kBinary:
      if (e.op == ",") return Print(*e.kids[0]) + ", " + Print(*e.kids[1]);
      return Print(*e.kids[0]) + " " + e.op + " " + Print(*e.kids[1]);
    case EK::kTernary:
      return Print(*e.kids[0]) + " ? " + Print(*e.kids[1]) + " : " +
             Print(*e.kids[2]);
    case EK::kCast:
      return "(" + e.type + ")" + Print(*e.kids[0]);
    case EK::kParen:
      return "(" + Print(*e.kids[0]) + ")";
    case EK::kSizeofType:
      return "sizeof(" + e.type + ")";
    case EK::kInitList: {
      std::string s = "{";
      for (size_t i = 0; i < e.kids.size(); ++i)
        s += (i ? ", " : "") + Print(*e.kids[i]);
      return s + "}";
    }
  }
  return "";
}

static std::string Ind(int n) { return std::string(2 * n, ' '); }

std::string PrintDecl(const Stmt &s) {
  std::string out = s.type + " ";
  for (size_t i = 0; i < s.decls.size(); ++i) {
    const Declarator &d = s.decls[i];
    out += (i ? ", " : "") + d.prefix + d.name + d.suffix;
    if (d.init && d.ctor_init)
      out += "(" + Print(*d.init) + ")";
    else if (d.init)
      out += " = " + Print(*d.init);
  }
  return out + ";";
}

std::string Print(const Stmt &s, int ind) {
  switch (s.kind) {
    case SK::kCompound: {
      std::string out = "{\n";
      for (const auto &k : s.kids) out += Print(*k, ind + 1);
      return out + Ind(ind) + "}\n";
    }
    case SK::kDecl:
      return Ind(ind) + PrintDecl(s) + "\n";
    case SK::kExpr:
      return Ind(ind) + Print(*s.expr) + ";\n";
    case SK::kIf: {
      std::string out = Ind(ind) + "if (" + Print(*s.expr) + ")";
      bool cth = s.kids[0]->kind == SK::kCompound;
      out += cth ? " " + Print(*s.kids[0], ind)
                 : "\n" + Print(*s.kids[0], ind + 1);
      if (s.kids.size() > 1) {
        bool cel = s.kids[1]->kind == SK::kCompound;
        bool elif = s.kids[1]->kind == SK::kIf;
        std::string el = cel ? " " + Print(*s.kids[1], ind)
                             : elif ? " " + Print(*s.kids[1], ind).substr(
                                               2 * ind)
                                    : "\n" + Print(*s.kids[1], ind + 1);
        out += Ind(ind) + "else" + el;
      }
      return out;
    }
    case SK::kFor: {
      std::string init = ";";
      if (s.init) {
        init = s.init->kind == SK::kDecl ? PrintDecl(*s.init)
                                         : Print(*s.init->expr) + ";";
      }
      std::string out = Ind(ind) + "for (" + init + " " +
                        (s.cond ? Print(*s.cond) : "") + "; " +
                        (s.step ? Print(*s.step) : "") + ") ";
      std::string b = s.kids[0]->kind == SK::kCompound
                          ? Print(*s.kids[0], ind)
                          : "\n" + Print(*s.kids[0], ind + 1);
      return out + b;
    }
    case SK::kWhile: {
      std::string b = s.kids[0]->kind == SK::kCompound
                          ? Print(*s.kids[0], ind)
                          : "\n" + Print(*s.kids[0], ind + 1);
      return Ind(ind) + "while (" + Print(*s.expr) + ") " + b;
    }
    case SK::kDo:
      return Ind(ind) + "do " + Print(*s.kids[0], ind) + Ind(ind) +
             "while (" + Print(*s.expr) + ");\n";
    case SK::kSwitch:
      return Ind(ind) + "switch (" + Print(*s.expr) + ") " +
             Print(*s.kids[0], ind);
    case SK::kCase:
      return Ind(ind) + "case " + Print(*s.expr) + ":\n";
    case SK::kDefault:
      return Ind(ind) + "default:\n";
    case SK::kReturn:
      return Ind(ind) + "return" + (s.expr ? " " + Print(*s.expr) : "") +
             ";\n";
    case SK::kBreak:
      return Ind(ind) + "break;\n";
    case SK::kContinue:
      return Ind(ind) + "continue;\n";
    case SK::kEmpty:
      return Ind(ind) + ";\n";
    case SK::kGoto:
      return Ind(ind) + "goto " + s.label + ";\n";
    case SK::kLabel:
      return s.label + ":\n";
  }
  return "";
}

std::string PrintTokens(const std::vector<Token> &toks) {
  std::string out;
  for (size_t i = 0; i < toks.size(); ++i) {
    if (i && (toks[i].space_before || toks[i].newline_before)) out += " ";
    out += toks[i].text;
  }
  return out;
}

}  // namespace physisc
