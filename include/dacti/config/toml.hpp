// dacti/config/toml.hpp -- reader for the TOML subset of the reference's case files.
//
// The reference reads its cases with toml++ (config/config.cpp:10-141,
// config/util.hpp:15-103).  This header parses what those files use: tables
// ([a.b]), arrays of tables ([[geometry]]), key = value pairs with bare or
// quoted keys (dotted keys too), basic and literal strings, integers,
// floats (exponents, signs, underscores, inf/nan), booleans, arrays (over
// several lines, trailing comma) and inline tables.  Dates and multi-line
// strings are not supported (no case uses them) and raise an error.
// Values keep their TOML type, so the strict typing of get_entry
// (an integer where a float is required is an error) can be applied.
#ifndef DACTI_TOML_HPP
#define DACTI_TOML_HPP

#include <cmath>
#include <cstdint>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace dacti::toml {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Value;
using Table = std::map<std::string, std::shared_ptr<Value>>;
using Array = std::vector<std::shared_ptr<Value>>;

struct Value {
    enum Kind { STRING, INTEGER, FLOAT, BOOLEAN, ARRAY, TABLE } kind = TABLE;
    std::string s;
    int64_t i = 0;
    double f = 0.0;
    bool b = false;
    Array a;
    Table t;
    bool array_of_tables = false;  // declared with [[...]]
    bool defined = false;          // a table header or inline table defined it

    bool is_table() const { return kind == TABLE; }
    bool is_array() const { return kind == ARRAY; }
    // child of a table (nullptr when absent or not a table)
    const Value *get(const std::string &key) const {
        if (kind != TABLE) return nullptr;
        auto it = t.find(key);
        return it == t.end() ? nullptr : it->second.get();
    }
    // dotted path lookup ("run.end_time")
    const Value *at_path(const std::string &path) const {
        const Value *v = this;
        size_t p = 0;
        while (v) {
            const size_t q = path.find('.', p);
            v = v->get(path.substr(p, q == std::string::npos ? std::string::npos : q - p));
            if (q == std::string::npos) break;
            p = q + 1;
        }
        return v;
    }
};

class Parser {
public:
    explicit Parser(std::string text) : src_(std::move(text)) {}

    std::shared_ptr<Value> parse() {
        auto root = std::make_shared<Value>();
        Value *cur = root.get();
        while (true) {
            skip_ws_nl();
            if (eof()) break;
            const char c = peek();
            if (c == '[') {
                const bool aot = peek(1) == '[';
                pos_ += aot ? 2 : 1;
                skip_ws();
                std::vector<std::string> keys = parse_key();
                skip_ws();
                expect(']');
                if (aot) expect(']');
                cur = open_table(root.get(), keys, aot);
            } else {
                parse_keyval(cur);
            }
            skip_ws();
            if (!eof() && peek() == '#') skip_comment();
            if (!eof() && peek() != '\n' && peek() != '\r') fail("expected a new line");
        }
        return root;
    }

private:
    std::string src_;
    size_t pos_ = 0;
    int line_ = 1;

    bool eof() const { return pos_ >= src_.size(); }
    char peek(size_t k = 0) const { return pos_ + k < src_.size() ? src_[pos_ + k] : '\0'; }
    [[noreturn]] void fail(const std::string &m) const {
        throw ParseError("TOML line " + std::to_string(line_) + ": " + m);
    }
    void expect(char c) {
        if (peek() != c) fail(std::string("expected '") + c + "'");
        ++pos_;
    }
    void skip_ws() {
        while (!eof() && (peek() == ' ' || peek() == '\t')) ++pos_;
    }
    void skip_comment() {
        while (!eof() && peek() != '\n') ++pos_;
    }
    void skip_ws_nl() {
        while (!eof()) {
            const char c = peek();
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

    std::string parse_simple_key() {
        if (peek() == '"') return parse_basic_string();
        if (peek() == '\'') return parse_literal_string();
        const size_t b = pos_;
        while (!eof()) {
            const char c = peek();
            if ((c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || (c >= '0' && c <= '9') || c == '_' || c == '-')
                ++pos_;
            else
                break;
        }
        if (pos_ == b) fail("expected a key");
        return src_.substr(b, pos_ - b);
    }

    std::vector<std::string> parse_key() {
        std::vector<std::string> keys{parse_simple_key()};
        while (true) {
            skip_ws();
            if (peek() != '.') break;
            ++pos_;
            skip_ws();
            keys.push_back(parse_simple_key());
        }
        return keys;
    }

    // table of a header [a.b.c] / [[a.b.c]]; intermediate tables are created
    // implicitly, an array of tables resolves to its last element
    Value *open_table(Value *root, const std::vector<std::string> &keys, bool aot) {
        Value *v = root;
        for (size_t k = 0; k < keys.size(); ++k) {
            const bool last = k + 1 == keys.size();
            auto it = v->t.find(keys[k]);
            if (it == v->t.end()) {
                auto nv = std::make_shared<Value>();
                if (last && aot) {
                    nv->kind = Value::ARRAY;
                    nv->array_of_tables = true;
                }
                it = v->t.emplace(keys[k], nv).first;
            }
            Value *c = it->second.get();
            if (last && aot) {
                if (!(c->kind == Value::ARRAY && c->array_of_tables)) fail("'" + keys[k] + "' is not an array of tables");
                auto e = std::make_shared<Value>();
                e->defined = true;
                c->a.push_back(e);
                return e.get();
            }
            if (c->kind == Value::ARRAY && c->array_of_tables) {
                v = c->a.back().get();
                continue;
            }
            if (c->kind != Value::TABLE) fail("'" + keys[k] + "' is not a table");
            if (last) {
                if (c->defined) fail("table '" + keys[k] + "' defined twice");
                c->defined = true;
            }
            v = c;
        }
        return v;
    }

    void parse_keyval(Value *table) {
        std::vector<std::string> keys = parse_key();
        skip_ws();
        expect('=');
        skip_ws();
        std::shared_ptr<Value> val = parse_value();
        Value *v = table;
        for (size_t k = 0; k + 1 < keys.size(); ++k) {
            auto it = v->t.find(keys[k]);
            if (it == v->t.end()) it = v->t.emplace(keys[k], std::make_shared<Value>()).first;
            if (it->second->kind != Value::TABLE) fail("'" + keys[k] + "' is not a table");
            v = it->second.get();
        }
        if (v->t.count(keys.back())) fail("key '" + keys.back() + "' defined twice");
        v->t.emplace(keys.back(), val);
    }

    std::string parse_basic_string() {
        expect('"');
        if (peek() == '"' && peek(1) == '"') fail("multi-line strings are not supported");
        std::string out;
        while (true) {
            if (eof() || peek() == '\n') fail("unterminated string");
            const char c = src_[pos_++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            const char e = src_[pos_++];
            switch (e) {
            case 'n': out += '\n'; break;
            case 't': out += '\t'; break;
            case 'r': out += '\r'; break;
            case '"': out += '"'; break;
            case '\\': out += '\\'; break;
            default: fail("unsupported escape");
            }
        }
        return out;
    }

    std::string parse_literal_string() {
        expect('\'');
        const size_t b = pos_;
        while (!eof() && peek() != '\'' && peek() != '\n') ++pos_;
        if (peek() != '\'') fail("unterminated string");
        std::string out = src_.substr(b, pos_ - b);
        ++pos_;
        return out;
    }

    std::shared_ptr<Value> parse_value() {
        auto v = std::make_shared<Value>();
        const char c = peek();
        if (c == '"') {
            v->kind = Value::STRING;
            v->s = parse_basic_string();
        } else if (c == '\'') {
            v->kind = Value::STRING;
            v->s = parse_literal_string();
        } else if (c == '[') {
            ++pos_;
            v->kind = Value::ARRAY;
            while (true) {
                skip_ws_nl();
                if (peek() == ']') {
                    ++pos_;
                    break;
                }
                v->a.push_back(parse_value());
                skip_ws_nl();
                if (peek() == ',') {
                    ++pos_;
                    continue;
                }
                skip_ws_nl();
                expect(']');
                break;
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
                    skip_ws();
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
        } else if (src_.compare(pos_, 4, "true") == 0) {
            v->kind = Value::BOOLEAN;
            v->b = true;
            pos_ += 4;
        } else if (src_.compare(pos_, 5, "false") == 0) {
            v->kind = Value::BOOLEAN;
            v->b = false;
            pos_ += 5;
        } else {
            parse_number(*v);
        }
        return v;
    }

    void parse_number(Value &v) {
        const size_t b = pos_;
        while (!eof()) {
            const char c = peek();
            if ((c >= '0' && c <= '9') || c == '+' || c == '-' || c == '.' || c == 'e' || c == 'E' || c == '_' ||
                c == 'i' || c == 'n' || c == 'f' || c == 'a')
                ++pos_;
            else
                break;
        }
        std::string tok;
        for (size_t k = b; k < pos_; ++k)
            if (src_[k] != '_') tok += src_[k];
        if (tok.empty()) fail("expected a value");
        std::string body = (tok[0] == '+' || tok[0] == '-') ? tok.substr(1) : tok;
        const double sign = tok[0] == '-' ? -1.0 : 1.0;
        if (body == "inf" || body == "nan") {
            v.kind = Value::FLOAT;
            v.f = body == "inf" ? sign * HUGE_VAL : std::nan("");
            return;
        }
        const bool is_float = body.find_first_of(".eE") != std::string::npos;
        if (body.size() > 1 && body[0] == '0' && !is_float) fail("leading zeros are not allowed");
        for (char ch : body)
            if (!((ch >= '0' && ch <= '9') || ch == '.' || ch == 'e' || ch == 'E' || ch == '+' || ch == '-'))
                fail("invalid number '" + tok + "'");
        size_t used = 0;
        try {
            if (is_float) {
                v.kind = Value::FLOAT;
                v.f = std::stod(tok, &used);
            } else {
                v.kind = Value::INTEGER;
                v.i = std::stoll(tok, &used);
            }
        } catch (const std::exception &) {
            fail("invalid number '" + tok + "'");
        }
        if (used != tok.size()) fail("invalid number '" + tok + "'");
    }
};

inline std::shared_ptr<Value> parse(const std::string &text) { return Parser(text).parse(); }

inline std::shared_ptr<Value> parse_file(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw ParseError("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    return parse(ss.str());
}

}  // namespace dacti::toml

#endif
