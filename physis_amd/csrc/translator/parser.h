// Parser and printer of the Physis C subset (see ast.h).
#ifndef PHYSIS_AMD_TRANSLATOR_PARSER_H_
#define PHYSIS_AMD_TRANSLATOR_PARSER_H_

#include <set>
#include <string>
#include <utility>
#include <vector>

#include "ast.h"

namespace physisc {

// PSGrid{1,2,3}D{Float,Double,Int,Long} and the grid types a program
// declares with DeclareGrid{1,2,3}D(Name, type): dimensionality and C element
// type ("float", ..., or the user type's text, e.g. "struct Point").
bool IsGridTypeName(const std::string &s, int *dims, std::string *elem);
// A grid type declared by DeclareGrid<dims>D(name, elem) (PSGrid<dims>D<name>).
void RegisterUserGridType(const std::string &name, int dims,
                          const std::string &elem);
bool IsUserGridType(const std::string &s);

// Members of a struct defined in the program, in declaration order: scalar
// base type, name and array extents (empty for a scalar member).
struct StructMember {
  std::string type;
  std::string name;
  std::vector<int> extents;
  int Components() const {
    int n = 1;
    for (int e : extents) n *= e;
    return n;
  }
};
// "struct Tag" (or a typedef name) -> members; empty if unknown.
const std::vector<StructMember> *FindStruct(const std::string &type);

class Parser {
 public:
  Parser(std::vector<Token> toks, std::string file);
  Program ParseProgram();
  // Global variables seen while parsing: (type text, name).
  const std::vector<std::pair<std::string, std::string>> &globals() const {
    return globals_;
  }

 private:
  [[noreturn]] void Fail(const std::string &msg) const;
  const Token &Peek(int k = 0) const;
  bool Accept(const char *s);
  void Expect(const char *s);
  bool IsTypeName(const std::string &s) const;
  static bool IsKeyword(const std::string &s);
  bool StartsDeclaration() const;
  bool LooksLikeCast() const;
  static int BinaryPrec(const std::string &op);

  std::string ParseBaseType();
  // struct body after '{' up to its '}' (registered under `tag`); returns
  // the body's text
  std::string ParseStructBody(const std::string &tag);
  std::string ParseTypeName();
  Declarator ParseDeclarator(bool allow_init);
  ExprP ParseInitList();
  ExprP ParseExpr();
  ExprP ParseAssign();
  ExprP ParseTernary();
  ExprP ParseBinary(int min_prec);
  ExprP ParseUnary();
  ExprP ParsePostfix();
  ExprP ParsePrimary();
  StmtP ParseStmt();
  StmtP ParseDecl();
  StmtP ParseCompound();

  std::vector<Token> t_;
  size_t pos_ = 0;
  std::string file_;
  std::set<std::string> typedefs_;
  std::vector<std::pair<std::string, std::string>> globals_;
};

std::string Print(const Expr &e);
std::string Print(const Stmt &s, int indent);
std::string PrintTokens(const std::vector<Token> &toks);

}  // namespace physisc

#endif  // PHYSIS_AMD_TRANSLATOR_PARSER_H_
