// toml.h -- the TOML subset the heddlemd input schema uses, parsed into an
// ordered tree: tables ([a.b]), arrays of tables ([[a]] with sub-tables
// [a.b] applying to the last element), key = value with dotted keys,
// basic and literal strings, integers, floats (exponents, underscores,
// inf/nan), booleans, arrays (multi-line, trailing commas) and inline
// tables.  Every array-of-tables element records its position in the
// document, which orders [[phase]] against [[minimization]].
#pragma once

#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace toml {

struct Value;
using Table = std::vector<std::pair<std::string, std::shared_ptr<Value>>>;

struct Value {
  enum Kind { STRING, INT, FLOAT, BOOL, ARRAY, TABLE } kind = TABLE;
  std::string s;
  int64_t i = 0;
  double f = 0.0;
  bool b = false;
  std::vector<std::shared_ptr<Value>> arr;
  Table tab;
  bool array_of_tables = false;  // an [[x]] array
  bool defined = false;          // a [x] header or key already defined this table
  size_t order = 0;              // document position of an [[x]] element
  int line = 0;

  const Value* get(const std::string& key) const {
    for (auto& kv : tab)
      if (kv.first == key) return kv.second.get();
    return nullptr;
  }
  Value* get_mut(const std::string& key) {
    for (auto& kv : tab)
      if (kv.first == key) return kv.second.get();
    return nullptr;
  }
  bool is_number() const { return kind == INT || kind == FLOAT; }
  double num() const { return kind == INT ? (double)i : f; }
};

struct ParseError : std::runtime_error {
  explicit ParseError(const std::string& m) : std::runtime_error(m) {}
};

class Parser {
 public:
  explicit Parser(const std::string& text) : t_(text) {}

  std::shared_ptr<Value> parse() {
    auto root = std::make_shared<Value>();
    root->kind = Value::TABLE;
    Value* cur = root.get();
    while (true) {
      skip_ws_comments_newlines();
      if (eof()) break;
      if (peek() == '[') {
        bool aot = peek(1) == '[';
        pos_ += aot ? 2 : 1;
        std::vector<std::string> path = parse_key_path();
        skip_ws();
        expect(']');
        if (aot) expect(']');
        end_of_line();
        cur = aot ? open_array_table(root.get(), path) : open_table(root.get(), path);
      } else {
        parse_keyval(cur);
        end_of_line();
      }
    }
    return root;
  }

 private:
  const std::string& t_;
  size_t pos_ = 0;
  int line_ = 1;
  size_t aot_counter_ = 0;

  bool eof() const { return pos_ >= t_.size(); }
  char peek(size_t k = 0) const { return pos_ + k < t_.size() ? t_[pos_ + k] : '\0'; }
  [[noreturn]] void fail(const std::string& m) const {
    throw ParseError("TOML parse error at line " + std::to_string(line_) + ": " + m);
  }
  void expect(char c) {
    if (peek() != c) fail(std::string("expected '") + c + "'");
    ++pos_;
  }
  void skip_ws() {
    while (!eof() && (peek() == ' ' || peek() == '\t')) ++pos_;
  }
  void skip_comment() {
    if (peek() == '#')
      while (!eof() && peek() != '\n') ++pos_;
  }
  void skip_ws_comments_newlines() {
    while (!eof()) {
      char c = peek();
      if (c == ' ' || c == '\t' || c == '\r') {
        ++pos_;
      } else if (c == '\n') {
        ++pos_;
        ++line_;
      } else if (c == '#') {
        skip_comment();
      } else {
        break;
      }
    }
  }
  void end_of_line() {
    skip_ws();
    skip_comment();
    if (peek() == '\r') ++pos_;
    if (!eof() && peek() != '\n') fail("unexpected text after a value");
  }

  std::string parse_simple_key() {
    skip_ws();
    if (peek() == '"' || peek() == '\'') return parse_string();
    size_t s = pos_;
    while (!eof() && (isalnum((unsigned char)peek()) || peek() == '_' || peek() == '-')) ++pos_;
    if (s == pos_) fail("expected a key");
    return t_.substr(s, pos_ - s);
  }
  std::vector<std::string> parse_key_path() {
    std::vector<std::string> p{parse_simple_key()};
    skip_ws();
    while (peek() == '.') {
      ++pos_;
      p.push_back(parse_simple_key());
      skip_ws();
    }
    return p;
  }

  Value* child_table(Value* t, const std::string& k, bool for_header) {
    Value* v = t->get_mut(k);
    if (!v) {
      auto nv = std::make_shared<Value>();
      nv->kind = Value::TABLE;
      nv->line = line_;
      t->tab.emplace_back(k, nv);
      return nv.get();
    }
    if (v->kind == Value::ARRAY && v->array_of_tables) {
      if (v->arr.empty()) fail("empty array of tables `" + k + "`");
      return v->arr.back().get();
    }
    if (v->kind != Value::TABLE) fail("key `" + k + "` is not a table");
    (void)for_header;
    return v;
  }
  Value* open_table(Value* root, const std::vector<std::string>& path) {
    Value* t = root;
    for (size_t i = 0; i + 1 < path.size(); ++i) t = child_table(t, path[i], true);
    Value* v = t->get_mut(path.back());
    if (v && v->kind == Value::TABLE && v->defined) fail("table [" + path.back() + "] defined twice");
    v = child_table(t, path.back(), true);
    v->defined = true;
    return v;
  }
  Value* open_array_table(Value* root, const std::vector<std::string>& path) {
    Value* t = root;
    for (size_t i = 0; i + 1 < path.size(); ++i) t = child_table(t, path[i], true);
    Value* a = t->get_mut(path.back());
    if (!a) {
      auto na = std::make_shared<Value>();
      na->kind = Value::ARRAY;
      na->array_of_tables = true;
      t->tab.emplace_back(path.back(), na);
      a = na.get();
    } else if (!(a->kind == Value::ARRAY && a->array_of_tables)) {
      fail("`" + path.back() + "` is not an array of tables");
    }
    auto e = std::make_shared<Value>();
    e->kind = Value::TABLE;
    e->order = aot_counter_++;
    e->line = line_;
    a->arr.push_back(e);
    return e.get();
  }

  void parse_keyval(Value* table) {
    std::vector<std::string> path = parse_key_path();
    skip_ws();
    expect('=');
    skip_ws();
    Value* t = table;
    for (size_t i = 0; i + 1 < path.size(); ++i) t = child_table(t, path[i], false);
    if (t->get(path.back())) fail("duplicate key `" + path.back() + "`");
    auto v = parse_value();
    v->line = line_;
    t->tab.emplace_back(path.back(), v);
  }

  std::string parse_string() {
    char q = peek();
    bool multi = peek(1) == q && peek(2) == q;
    pos_ += multi ? 3 : 1;
    std::string out;
    if (multi && peek() == '\n') {
      ++pos_;
      ++line_;
    }
    while (true) {
      if (eof()) fail("unterminated string");
      char c = peek();
      if (multi ? (c == q && peek(1) == q && peek(2) == q) : c == q) {
        pos_ += multi ? 3 : 1;
        break;
      }
      if (c == '\n') {
        if (!multi) fail("newline in a string");
        ++line_;
      }
      if (c == '\\' && q == '"') {
        ++pos_;
        char e = peek();
        ++pos_;
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case '"': out += '"'; break;
          case '\\': out += '\\'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u':
          case 'U': {
            int nd = e == 'u' ? 4 : 8;
            uint32_t cp = (uint32_t)std::stoul(t_.substr(pos_, nd), nullptr, 16);
            pos_ += nd;
            if (cp < 0x80) {
              out += (char)cp;
            } else if (cp < 0x800) {
              out += (char)(0xC0 | (cp >> 6));
              out += (char)(0x80 | (cp & 0x3F));
            } else if (cp < 0x10000) {
              out += (char)(0xE0 | (cp >> 12));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            } else {
              out += (char)(0xF0 | (cp >> 18));
              out += (char)(0x80 | ((cp >> 12) & 0x3F));
              out += (char)(0x80 | ((cp >> 6) & 0x3F));
              out += (char)(0x80 | (cp & 0x3F));
            }
            break;
          }
          default: fail("unknown escape");
        }
        continue;
      }
      out += c;
      ++pos_;
    }
    return out;
  }

  std::shared_ptr<Value> parse_value() {
    auto v = std::make_shared<Value>();
    char c = peek();
    if (c == '"' || c == '\'') {
      v->kind = Value::STRING;
      v->s = parse_string();
    } else if (c == '[') {
      ++pos_;
      v->kind = Value::ARRAY;
      while (true) {
        skip_ws_comments_newlines();
        if (peek() == ']') {
          ++pos_;
          break;
        }
        v->arr.push_back(parse_value());
        skip_ws_comments_newlines();
        if (peek() == ',') {
          ++pos_;
        } else if (peek() == ']') {
          ++pos_;
          break;
        } else {
          fail("expected ',' or ']' in an array");
        }
      }
    } else if (c == '{') {
      ++pos_;
      v->kind = Value::TABLE;
      v->defined = true;
      skip_ws();
      if (peek() == '}') {
        ++pos_;
      } else {
        while (true) {
          parse_keyval(v.get());
          skip_ws();
          if (peek() == ',') {
            ++pos_;
            continue;
          }
          expect('}');
          break;
        }
      }
    } else if (t_.compare(pos_, 4, "true") == 0) {
      pos_ += 4;
      v->kind = Value::BOOL;
      v->b = true;
    } else if (t_.compare(pos_, 5, "false") == 0) {
      pos_ += 5;
      v->kind = Value::BOOL;
      v->b = false;
    } else {
      parse_number(*v);
    }
    return v;
  }

  void parse_number(Value& v) {
    size_t s = pos_;
    while (!eof()) {
      char c = peek();
      if (isalnum((unsigned char)c) || c == '+' || c == '-' || c == '.' || c == '_')
        ++pos_;
      else
        break;
    }
    std::string raw = t_.substr(s, pos_ - s), num;
    for (char c : raw)
      if (c != '_') num += c;
    if (num.empty()) fail("expected a value");
    std::string body = (num[0] == '+' || num[0] == '-') ? num.substr(1) : num;
    if (body == "inf" || body == "nan") {
      v.kind = Value::FLOAT;
      double x = body == "inf" ? HUGE_VAL : NAN;
      v.f = num[0] == '-' ? -x : x;
      return;
    }
    bool is_float = num.find_first_of(".eE") != std::string::npos &&
                    !(num.size() > 2 && num[0] == '0' && (num[1] == 'x' || num[1] == 'o' || num[1] == 'b'));
    char* end = nullptr;
    if (is_float) {
      v.kind = Value::FLOAT;
      v.f = strtod(num.c_str(), &end);
    } else {
      v.kind = Value::INT;
      int base = 10;
      std::string digits = num;
      if (body.size() > 2 && body[0] == '0' && (body[1] == 'x' || body[1] == 'o' || body[1] == 'b')) {
        base = body[1] == 'x' ? 16 : body[1] == 'o' ? 8 : 2;
        digits = (num[0] == '-' ? "-" : "") + body.substr(2);
      }
      v.i = strtoll(digits.c_str(), &end, base);
    }
    if (!end || *end != '\0') fail("malformed number `" + raw + "`");
  }
};

inline std::shared_ptr<Value> parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace toml
