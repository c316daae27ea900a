// Lexer for the Physis C subset (after preprocessing).
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "ast.h"

namespace physisc {

namespace {

const char *kPuncts[] = {
    "<<=", ">>=", "...", "->", "++", "--", "<<", ">>", "<=", ">=", "==", "!=",
    "&&", "||", "*=", "/=", "%=", "+=", "-=", "&=", "^=", "|=", "##",
    "[", "]", "(", ")", "{", "}", ".", "&", "*", "+", "-", "~", "!", "/",
    "%", "<", ">", "^", "|", "?", ":", ";", "=", ",", "#"};

[[noreturn]] void Fail(const std::string &file, int line, const char *msg) {
  std::fprintf(stderr, "%s:%d: physisc-hip: %s\n", file.c_str(), line, msg);
  std::exit(1);
}

}  // namespace

std::vector<Token> Lex(const std::string &src, const std::string &file) {
  std::vector<Token> out;
  size_t i = 0, n = src.size();
  int line = 1;
  bool space = false, nl = true;
  while (i < n) {
    char c = src[i];
    if (c == '\n') {
      ++line;
      ++i;
      space = nl = true;
      continue;
    }
    if (std::isspace((unsigned char)c)) {
      ++i;
      space = true;
      continue;
    }
    if (c == '/' && i + 1 < n && src[i + 1] == '/') {
      while (i < n && src[i] != '\n') ++i;
      space = true;
      continue;
    }
    if (c == '/' && i + 1 < n && src[i + 1] == '*') {
      i += 2;
      while (i + 1 < n && !(src[i] == '*' && src[i + 1] == '/')) {
        if (src[i] == '\n') ++line;
        ++i;
      }
      i += 2;
      space = true;
      continue;
    }
    // Preprocessor leftovers (line markers, #pragma) are skipped whole.
    if (c == '#' && nl) {
      while (i < n && src[i] != '\n') ++i;
      continue;
    }
    Token t;
    t.line = line;
    t.space_before = space;
    t.newline_before = nl;
    space = nl = false;
    if (std::isalpha((unsigned char)c) || c == '_') {
      size_t j = i;
      while (j < n && (std::isalnum((unsigned char)src[j]) || src[j] == '_')) ++j;
      t.kind = Tok::kIdent;
      t.text = src.substr(i, j - i);
      i = j;
    } else if (std::isdigit((unsigned char)c) ||
               (c == '.' && i + 1 < n && std::isdigit((unsigned char)src[i + 1]))) {
      size_t j = i;
      while (j < n) {
        char d = src[j];
        if (std::isalnum((unsigned char)d) || d == '.' || d == '_') {
          ++j;
        } else if ((d == '+' || d == '-') &&
                   (src[j - 1] == 'e' || src[j - 1] == 'E' ||
                    src[j - 1] == 'p' || src[j - 1] == 'P') &&
                   !(src[i] == '0' && j > i + 1 &&
                     (src[i + 1] == 'x' || src[i + 1] == 'X') &&
                     (src[j - 1] == 'e' || src[j - 1] == 'E'))) {
          ++j;
        } else {
          break;
        }
      }
      t.kind = Tok::kNumber;
      t.text = src.substr(i, j - i);
      i = j;
    } else if (c == '"' || c == '\'') {
      size_t j = i + 1;
      while (j < n && src[j] != c) {
        if (src[j] == '\\') ++j;
        if (src[j] == '\n') Fail(file, line, "newline in literal");
        ++j;
      }
      if (j >= n) Fail(file, line, "unterminated literal");
      t.kind = c == '"' ? Tok::kString : Tok::kChar;
      t.text = src.substr(i, j + 1 - i);
      i = j + 1;
    } else {
      const char *match = nullptr;
      for (const char *p : kPuncts) {
        size_t len = std::strlen(p);
        if (src.compare(i, len, p) == 0) {
          match = p;
          break;
        }
      }
      if (!match) Fail(file, line, "unexpected character");
      t.kind = Tok::kPunct;
      t.text = match;
      i += t.text.size();
    }
    out.push_back(t);
  }
  Token end;
  end.kind = Tok::kEnd;
  end.line = line;
  out.push_back(end);
  return out;
}

}  // namespace physisc
