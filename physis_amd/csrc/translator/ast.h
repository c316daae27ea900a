// Tokens and syntax tree of the Physis C subset accepted by physisc-hip.
//
// The reference translator works on a ROSE AST (translator/translation_
// context.cc:439-490); ROSE is not available, so physisc-hip carries its own
// small tree: enough C to hold stencil functions (expressions, declarations,
// if/for/while/do/switch/return) and to find and rewrite the PS* intrinsics
// in host code.  Host code outside intrinsics is kept as token runs, and so
// are top-level C++ declarations (class, using, namespace, template).
#ifndef PHYSIS_AMD_TRANSLATOR_AST_H_
#define PHYSIS_AMD_TRANSLATOR_AST_H_

#include <memory>
#include <string>
#include <vector>

namespace physisc {

enum class Tok { kIdent, kNumber, kString, kChar, kPunct, kEnd };

struct Token {
  Tok kind = Tok::kEnd;
  std::string text;
  int line = 0;
  bool space_before = false;   // whitespace preceded the token
  bool newline_before = false; // a line break preceded the token
  bool Is(const char *s) const { return text == s && kind != Tok::kString; }
};

std::vector<Token> Lex(const std::string &src, const std::string &file);

// ------------------------------------------------------------ expressions
struct Expr;
typedef std::unique_ptr<Expr> ExprP;

enum class EK {
  kIdent,     // text
  kLiteral,   // text (number, string, char)
  kCall,      // kids[0] = callee, kids[1..] = args
  kIndex,     // kids[0][kids[1]]
  kMember,    // kids[0] op text   (op "." or "->")
  kUnary,     // op kids[0]        (prefix: - + ! ~ * & ++ --)
  kPostfix,   // kids[0] op        (++ --)
  kBinary,    // kids[0] op kids[1] (includes assignments and comma)
  kTernary,   // kids[0] ? kids[1] : kids[2]
  kCast,      // (type) kids[0]
  kParen,     // ( kids[0] )
  kSizeofType,// sizeof(type)
  kInitList,  // { kids... }
  kRaw,       // text emitted verbatim (produced by rewriting)
};

struct Expr {
  EK kind;
  std::string op;    // operator or member operator
  std::string text;  // identifier / literal / member name / raw text
  std::string type;  // cast / sizeof type, as source text
  std::vector<ExprP> kids;
  int line = 0;
  explicit Expr(EK k) : kind(k) {}
};

// ------------------------------------------------------------- statements
struct Stmt;
typedef std::unique_ptr<Stmt> StmtP;

struct Declarator {
  std::string name;
  std::string prefix;  // pointer stars / qualifiers before the name
  std::string suffix;  // array bounds after the name
  ExprP init;
  bool ctor_init = false;  // C++ "T v(args)": init holds the argument list
};

enum class SK {
  kCompound, kDecl, kExpr, kIf, kFor, kWhile, kDo, kReturn, kBreak,
  kContinue, kSwitch, kCase, kDefault, kEmpty, kGoto, kLabel
};

struct Stmt {
  SK kind;
  // kDecl: base type text + declarators
  std::string type;
  std::vector<Declarator> decls;
  // kExpr / kReturn / kCase / kIf cond / kWhile cond / kSwitch cond
  ExprP expr;
  // kFor: init statement (decl or expr), cond, step
  StmtP init;
  ExprP cond, step;
  // bodies: kCompound children, kIf then/else, loops body
  std::vector<StmtP> kids;
  std::string label;
  int line = 0;
  explicit Stmt(SK k) : kind(k) {}
};

// ---------------------------------------------------------- top level
struct Param {
  std::string type;  // full type text (base + pointer stars)
  std::string name;
};

struct Function {
  std::string ret_type;      // return type text (without storage class)
  bool is_static = false;
  bool is_inline = false;
  std::string name;
  std::vector<Param> params;
  bool variadic = false;
  StmtP body;                // null for a prototype
  int line = 0;
};

// A top-level item: a function or any other declaration kept as tokens.
struct TopItem {
  enum Kind { kFunction, kTokens, kInclude } kind;
  std::unique_ptr<Function> func;
  std::vector<Token> tokens;   // kTokens: verbatim declaration
  std::string include_line;    // kInclude
  std::string user_grid;       // kTokens of DeclareGrid<n>D(Name, type): Name
  bool cxx = false;            // kTokens of a C++-only declaration (class,
                               // using, namespace, template): no C linkage
};

struct Program {
  std::string file;
  std::vector<TopItem> items;
};

}  // namespace physisc

#endif  // PHYSIS_AMD_TRANSLATOR_AST_H_
