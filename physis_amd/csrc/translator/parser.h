// Parser and printer of the Physis C subset (see ast.h).
#ifndef PHYSIS_AMD_TRANSLATOR_PARSER_H_
#define PHYSIS_AMD_TRANSLATOR_PARSER_H_

#include <set>
#include <string>
#include <utility>
#include <vector>

#include "ast.h"

namespace physisc {

// PSGrid{1,2,3}D{Float,Double,Int,Long}: dimensionality and C element type.
bool IsGridTypeName(const std::string &s, int *dims, std::string *elem);

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
