// Broombridge reader and IngestBroombridge restatement; see broombridge.hpp.
#include "broombridge.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdlib>
#include <limits>

#include "combinadic.hpp"

namespace qpe {

namespace {

[[noreturn]] void fail(int line, const std::string &msg) {
    throw FormatError(line > 0 ? "line " + std::to_string(line) + ": " + msg : msg);
}

struct Line {
    int indent;
    std::string text;
    int lineno;
};

bool is_space(char c) { return c == ' ' || c == '\t'; }

// Strips a "#" comment (one that starts a line or follows whitespace, outside
// quotes) and trailing whitespace.
std::string strip_comment(const std::string &s, int lineno) {
    char quote = 0;
    for (size_t k = 0; k < s.size(); ++k) {
        char c = s[k];
        if (quote == '\'') {
            if (c == '\'') {
                if (k + 1 < s.size() && s[k + 1] == '\'') ++k;
                else quote = 0;
            }
        } else if (quote == '"') {
            if (c == '\\') ++k;
            else if (c == '"') quote = 0;
        } else if (c == '#' && (k == 0 || is_space(s[k - 1]))) {
            std::string t = s.substr(0, k);
            while (!t.empty() && is_space(t.back())) t.pop_back();
            return t;
        } else if ((c == '\'' || c == '"') && (k == 0 || !std::isalnum((unsigned char)s[k - 1]))) {
            quote = c;
        }
    }
    if (quote) fail(lineno, "quoted scalar continues past the end of the line (not supported)");
    std::string t = s;
    while (!t.empty() && is_space(t.back())) t.pop_back();
    return t;
}

bool is_seq_item(const std::string &t) { return !t.empty() && t[0] == '-' && (t.size() == 1 || t[1] == ' '); }

// Position just past a quoted scalar that starts at s[k].
size_t skip_quoted(const std::string &s, size_t k, int lineno) {
    char q = s[k];
    for (size_t j = k + 1; j < s.size(); ++j) {
        if (q == '\'' && s[j] == '\'') {
            if (j + 1 < s.size() && s[j + 1] == '\'') { ++j; continue; }
            return j + 1;
        }
        if (q == '"') {
            if (s[j] == '\\') { ++j; continue; }
            if (s[j] == '"') return j + 1;
        }
    }
    fail(lineno, "unterminated quoted scalar");
}

std::string unquote(const std::string &s, int lineno) {
    std::string out;
    if (s[0] == '\'') {
        for (size_t j = 1; j + 1 < s.size(); ++j) {
            out += s[j];
            if (s[j] == '\'') ++j;  // '' -> '
        }
        return out;
    }
    for (size_t j = 1; j + 1 < s.size(); ++j) {
        char c = s[j];
        if (c != '\\') { out += c; continue; }
        char e = s[++j];
        switch (e) {
        case 'n': out += '\n'; break;
        case 't': out += '\t'; break;
        case '\\': out += '\\'; break;
        case '"': out += '"'; break;
        case '/': out += '/'; break;
        case '0': out += '\0'; break;
        default: fail(lineno, std::string("unsupported escape \\") + e);
        }
    }
    return out;
}

YamlNode scalar_node(const std::string &raw, int lineno) {
    YamlNode n;
    n.line = lineno;
    if (raw.empty() || raw == "~" || raw == "null" || raw == "Null" || raw == "NULL") return n;
    char c = raw[0];
    if (c == '&' || c == '*' || c == '!') fail(lineno, "anchors, aliases and tags are not supported");
    if (c == '|' || c == '>') fail(lineno, "block scalars are not supported");
    n.kind = YamlNode::kScalar;
    if (c == '\'' || c == '"') {
        if (skip_quoted(raw, 0, lineno) != raw.size()) fail(lineno, "text after a quoted scalar");
        n.scalar = unquote(raw, lineno);
    } else {
        n.scalar = raw;
    }
    return n;
}

// Splits "key: rest" at the first ": " (or a final ":") outside quotes.
// Returns false when the line is not a mapping entry.
bool split_key(const std::string &t, std::string &key, std::string &rest, int lineno) {
    if (t.empty() || t[0] == '[' || t[0] == '{') return false;
    size_t k = 0;
    if (t[0] == '\'' || t[0] == '"') {
        k = skip_quoted(t, 0, lineno);
        size_t j = k;
        while (j < t.size() && is_space(t[j])) ++j;
        if (j >= t.size() || t[j] != ':' || (j + 1 < t.size() && !is_space(t[j + 1]))) return false;
        key = unquote(t.substr(0, k), lineno);
        k = j;
    } else {
        for (;; ++k) {
            if (k >= t.size()) return false;
            if (t[k] == ':' && (k + 1 == t.size() || is_space(t[k + 1]))) break;
        }
        key = t.substr(0, k);
        while (!key.empty() && is_space(key.back())) key.pop_back();
    }
    size_t j = k + 1;
    while (j < t.size() && is_space(t[j])) ++j;
    rest = t.substr(j);
    return true;
}

class FlowParser {
  public:
    FlowParser(const std::string &s, int lineno) : s_(s), line_(lineno) {}

    YamlNode parse_all() {
        YamlNode n = value();
        ws();
        if (k_ != s_.size()) fail(line_, "text after a flow collection");
        return n;
    }

  private:
    const std::string &s_;
    size_t k_ = 0;
    int line_;

    void ws() {
        while (k_ < s_.size() && (is_space(s_[k_]) || s_[k_] == '\n')) ++k_;
    }

    // a scalar inside a flow collection; keys stop at ": ", values at , ] }
    std::string scalar_text(bool key) {
        ws();
        if (k_ < s_.size() && (s_[k_] == '\'' || s_[k_] == '"')) {
            size_t e = skip_quoted(s_, k_, line_);
            std::string q = s_.substr(k_, e - k_);
            k_ = e;
            return q;
        }
        size_t b = k_;
        while (k_ < s_.size()) {
            char c = s_[k_];
            if (c == ',' || c == ']' || c == '}') break;
            if (key && c == ':' && (k_ + 1 == s_.size() || is_space(s_[k_ + 1]) || s_[k_ + 1] == ',' ||
                                    s_[k_ + 1] == '}' || s_[k_ + 1] == ']'))
                break;
            if (c == '[' || c == '{') fail(line_, "unexpected bracket inside a plain scalar");
            ++k_;
        }
        std::string t = s_.substr(b, k_ - b);
        while (!t.empty() && (is_space(t.back()) || t.back() == '\n')) t.pop_back();
        return t;
    }

    YamlNode value() {
        ws();
        if (k_ >= s_.size()) fail(line_, "flow collection ends early");
        char c = s_[k_];
        if (c == '[') {
            ++k_;
            YamlNode n;
            n.kind = YamlNode::kSeq;
            n.line = line_;
            for (;;) {
                ws();
                if (k_ >= s_.size()) fail(line_, "unterminated flow sequence");
                if (s_[k_] == ']') { ++k_; break; }
                n.seq.push_back(value());
                ws();
                if (k_ < s_.size() && s_[k_] == ',') { ++k_; continue; }
                if (k_ < s_.size() && s_[k_] == ']') { ++k_; break; }
                fail(line_, "expected ',' or ']' in a flow sequence");
            }
            return n;
        }
        if (c == '{') {
            ++k_;
            YamlNode n;
            n.kind = YamlNode::kMap;
            n.line = line_;
            for (;;) {
                ws();
                if (k_ >= s_.size()) fail(line_, "unterminated flow mapping");
                if (s_[k_] == '}') { ++k_; break; }
                std::string kt = scalar_text(true);
                std::string key = (!kt.empty() && (kt[0] == '\'' || kt[0] == '"')) ? unquote(kt, line_) : kt;
                ws();
                if (k_ >= s_.size() || s_[k_] != ':') fail(line_, "expected ':' after a flow mapping key");
                ++k_;
                ws();
                YamlNode v;
                v.line = line_;
                if (k_ < s_.size() && s_[k_] != ',' && s_[k_] != '}') v = value();
                n.map.emplace_back(key, std::move(v));
                ws();
                if (k_ < s_.size() && s_[k_] == ',') { ++k_; continue; }
                if (k_ < s_.size() && s_[k_] == '}') { ++k_; break; }
                fail(line_, "expected ',' or '}' in a flow mapping");
            }
            return n;
        }
        return scalar_node(scalar_text(false), line_);
    }
};

class BlockParser {
  public:
    explicit BlockParser(const std::string &text) {
        int lineno = 0;
        size_t b = 0;
        while (b <= text.size()) {
            size_t e = text.find('\n', b);
            if (e == std::string::npos) e = text.size();
            std::string raw = text.substr(b, e - b);
            b = e + 1;
            ++lineno;
            if (!raw.empty() && raw.back() == '\r') raw.pop_back();
            size_t ind = 0;
            while (ind < raw.size() && raw[ind] == ' ') ++ind;
            if (ind < raw.size() && raw[ind] == '\t') fail(lineno, "tab in indentation");
            std::string t = strip_comment(raw.substr(ind), lineno);
            if (t.empty()) continue;
            if (ind == 0 && (t == "---" || t.rfind("--- ", 0) == 0 || t[0] == '%')) {
                if (t.size() > 4 && t[0] == '-') t = t.substr(4);
                else continue;
            }
            if (ind == 0 && t == "...") break;
            lines_.push_back({(int)ind, t, lineno});
            if (e == text.size()) break;
        }
    }

    YamlNode parse_document() {
        if (lines_.empty()) return YamlNode();
        YamlNode n = node(lines_[0].indent, -1);
        if (i_ < lines_.size()) fail(lines_[i_].lineno, "unexpected indentation");
        return n;
    }

  private:
    std::vector<Line> lines_;
    size_t i_ = 0;

    // The block node whose first line is lines_[i_] (at indent ind);
    // parent is the indent of the enclosing collection.
    YamlNode node(int ind, int parent) {
        const std::string &t = lines_[i_].text;
        std::string key, rest;
        if (is_seq_item(t)) return sequence(ind);
        if (split_key(t, key, rest, lines_[i_].lineno)) return mapping(ind);
        return inline_value(t, parent);
    }

    YamlNode sequence(int ind) {
        YamlNode n;
        n.kind = YamlNode::kSeq;
        n.line = lines_[i_].lineno;
        while (i_ < lines_.size() && lines_[i_].indent == ind && is_seq_item(lines_[i_].text)) {
            Line &ln = lines_[i_];
            size_t k = 1;
            while (k < ln.text.size() && ln.text[k] == ' ') ++k;
            if (k >= ln.text.size()) {
                int at = ln.lineno;
                ++i_;
                if (i_ < lines_.size() && lines_[i_].indent > ind) {
                    n.seq.push_back(node(lines_[i_].indent, ind));
                } else {
                    YamlNode empty;
                    empty.line = at;
                    n.seq.push_back(empty);
                }
            } else {
                ln.indent = ind + (int)k;  // the item's content starts a node at its own column
                ln.text = ln.text.substr(k);
                n.seq.push_back(node(ln.indent, ind));
            }
            if (i_ < lines_.size() && lines_[i_].indent > ind) fail(lines_[i_].lineno, "unexpected indentation");
        }
        return n;
    }

    YamlNode mapping(int ind) {
        YamlNode n;
        n.kind = YamlNode::kMap;
        n.line = lines_[i_].lineno;
        while (i_ < lines_.size() && lines_[i_].indent == ind) {
            const int lineno = lines_[i_].lineno;
            std::string key, rest;
            if (!split_key(lines_[i_].text, key, rest, lineno)) fail(lineno, "expected 'key: value'");
            YamlNode v;
            v.line = lineno;
            if (rest.empty()) {
                ++i_;
                if (i_ < lines_.size() && lines_[i_].indent > ind)
                    v = node(lines_[i_].indent, ind);
                else if (i_ < lines_.size() && lines_[i_].indent == ind && is_seq_item(lines_[i_].text))
                    v = sequence(ind);  // "key:\n- item" at the key's own indent
            } else {
                v = inline_value(rest, ind);
            }
            n.map.emplace_back(key, std::move(v));
            if (i_ < lines_.size() && lines_[i_].indent > ind) fail(lines_[i_].lineno, "unexpected indentation");
        }
        return n;
    }

    // A value written on the current line: a flow collection (which may span
    // lines) or a scalar (plain scalars continue on lines indented past parent).
    YamlNode inline_value(const std::string &text, int parent) {
        const int lineno = lines_[i_].lineno;
        if (text[0] == '[' || text[0] == '{') {
            std::string buf = text;
            while (depth(buf, lineno) > 0) {
                ++i_;
                if (i_ >= lines_.size()) fail(lineno, "unterminated flow collection");
                buf += '\n';
                buf += lines_[i_].text;
            }
            ++i_;
            return FlowParser(buf, lineno).parse_all();
        }
        std::string buf = text;
        ++i_;
        if (text[0] != '\'' && text[0] != '"') {
            while (i_ < lines_.size() && lines_[i_].indent > parent && !is_seq_item(lines_[i_].text)) {
                std::string k2, r2;
                if (split_key(lines_[i_].text, k2, r2, lines_[i_].lineno)) break;
                buf += ' ';
                buf += lines_[i_].text;
                ++i_;
            }
        }
        return scalar_node(buf, lineno);
    }

    static int depth(const std::string &s, int lineno) {
        int d = 0;
        for (size_t k = 0; k < s.size(); ++k) {
            char c = s[k];
            if (c == '\'' || c == '"') {
                if (k == 0 || !std::isalnum((unsigned char)s[k - 1])) { k = skip_quoted(s, k, lineno) - 1; continue; }
            }
            if (c == '[' || c == '{') ++d;
            if (c == ']' || c == '}') --d;
        }
        return d;
    }
};

int popcount64(uint64_t x) { return __builtin_popcountll(x); }

// One "(<orbital><a|b>)" or "(<orbital><a|b>)+" operator (ParsePolishNotation,
// cpp/hamiltonian.cc:146-158): 1-based orbital, spin a (up) or b (down),
// "+" for creation.
struct LadderOp {
    unsigned orbital;  // 0-based
    bool up;
    bool raising;
};

LadderOp parse_ladder(const YamlNode &node) {
    if (node.kind != YamlNode::kScalar) fail(node.line, "expected a ladder operator such as (1a)+");
    const std::string &s = node.scalar;
    size_t k = 0;
    auto bad = [&]() { fail(node.line, "'" + s + "' is not a ladder operator of the form (<orbital><a|b>) or (<orbital><a|b>)+"); };
    if (s.size() < 4 || s[k++] != '(') bad();
    size_t b = k;
    while (k < s.size() && std::isdigit((unsigned char)s[k])) ++k;
    if (k == b || k - b > 6) bad();
    unsigned orb = (unsigned)std::stoul(s.substr(b, k - b));
    if (k >= s.size() || (s[k] != 'a' && s[k] != 'b')) bad();
    bool up = s[k++] == 'a';
    if (k >= s.size() || s[k++] != ')') bad();
    bool raising = false;
    if (k < s.size() && s[k] == '+') { raising = true; ++k; }
    if (k != s.size()) bad();
    if (orb < 1) fail(node.line, "Broombridge orbitals are 1-based");
    return {orb - 1, up, raising};
}

}  // namespace

// ------------------------------------------------------------------ YamlNode

const YamlNode *YamlNode::find(const std::string &key) const {
    if (kind != kMap) return nullptr;
    for (const auto &kv : map)
        if (kv.first == key) return &kv.second;
    return nullptr;
}

const YamlNode &YamlNode::at(const std::string &key) const {
    if (kind != kMap) fail(line, "expected a mapping holding '" + key + "'");
    const YamlNode *n = find(key);
    if (!n) fail(line, "missing key '" + key + "'");
    return *n;
}

const YamlNode &YamlNode::at(size_t i) const {
    if (kind != kSeq) fail(line, "expected a sequence");
    if (i >= seq.size()) fail(line, "sequence has " + std::to_string(seq.size()) + " entries, needed " +
                                        std::to_string(i + 1));
    return seq[i];
}

double YamlNode::as_double() const {
    if (kind != kScalar) fail(line, "expected a number");
    const std::string &s = scalar;
    if (s == ".inf" || s == "+.inf" || s == ".Inf" || s == "+.Inf") return std::numeric_limits<double>::infinity();
    if (s == "-.inf" || s == "-.Inf") return -std::numeric_limits<double>::infinity();
    if (s == ".nan" || s == ".NaN") return std::numeric_limits<double>::quiet_NaN();
    char *end = nullptr;
    errno = 0;
    double v = std::strtod(s.c_str(), &end);
    bool hexish = s.find_first_of("xXpP") != std::string::npos;
    if (s.empty() || end != s.c_str() + s.size() || hexish || std::isalpha((unsigned char)s[0]))
        fail(line, "'" + s + "' is not a number");
    return v;
}

unsigned YamlNode::as_unsigned() const {
    if (kind != kScalar) fail(line, "expected an unsigned integer");
    const std::string &s = scalar;
    size_t k = (!s.empty() && s[0] == '+') ? 1 : 0;
    if (k >= s.size() || s.size() - k > 9) fail(line, "'" + s + "' is not an unsigned integer");
    for (size_t j = k; j < s.size(); ++j)
        if (!std::isdigit((unsigned char)s[j])) fail(line, "'" + s + "' is not an unsigned integer");
    return (unsigned)std::stoul(s.substr(k));
}

YamlNode parse_yaml(const std::string &text) { return BlockParser(text).parse_document(); }

// ------------------------------------------------------------------ ingest

// cpp/hamiltonian.cc:8-144 without the StateVector construction (:130-141),
// which the caller does with the returned counts and determinants.
BroombridgeData read_broombridge(const std::string &text, unsigned prior_orbitals) {
    const YamlNode root = parse_yaml(text);
    BroombridgeData d;
    const YamlNode &integrals = root.at("integral_sets").at(0);
    const YamlNode &ham = integrals.at("hamiltonian");
    // :15 identity = coulomb_repulsion + energy_offset
    d.identity = integrals.at("coulomb_repulsion").at("value").as_double() +
                 integrals.at("energy_offset").at("value").as_double();
    if (!std::isfinite(d.identity)) fail(integrals.line, "non-finite coulomb_repulsion or energy_offset");

    unsigned maxorb = prior_orbitals;
    // :22-37 one-body [p, q, v], swapped so that p >= q
    const YamlNode &one = ham.at("one_electron_integrals").at("values");
    if (one.kind != YamlNode::kSeq && one.kind != YamlNode::kNull) fail(one.line, "one-body values must be a sequence");
    for (const YamlNode &t : one.seq) {
        unsigned p = t.at(0).as_unsigned(), q = t.at(1).as_unsigned();
        double v = t.at(2).as_double();
        if (p < 1 || q < 1) fail(t.line, "Broombridge orbitals are 1-based");
        if (!std::isfinite(v)) fail(t.line, "non-finite integral");
        if (q > p) std::swap(p, q);
        d.one_body.push_back({{p, q}, v});
        maxorb = std::max(maxorb, p);
    }
    // :39-52 two-body [i, l, j, k, v] (Mulliken) -> AddTwoBodyTerm(i, j, k, l)
    const YamlNode &two = ham.at("two_electron_integrals").at("values");
    if (two.kind != YamlNode::kSeq && two.kind != YamlNode::kNull) fail(two.line, "two-body values must be a sequence");
    for (const YamlNode &t : two.seq) {
        unsigned i = t.at(0).as_unsigned(), l = t.at(1).as_unsigned();
        unsigned j = t.at(2).as_unsigned(), k = t.at(3).as_unsigned();
        double v = t.at(4).as_double();
        if (i < 1 || j < 1 || k < 1 || l < 1) fail(t.line, "Broombridge orbitals are 1-based");
        if (!std::isfinite(v)) fail(t.line, "non-finite integral");
        d.two_body.push_back({{i, j, k, l}, v});
        maxorb = std::max(maxorb, std::max(std::max(i, j), std::max(k, l)));
    }
    if (maxorb < 1 || maxorb > 32) fail(ham.line, "orbital count " + std::to_string(maxorb) + " outside [1, 32]");
    d.num_orbitals = maxorb;

    // :55-128 the first suggested state: [coefficient, op, op, ..., last]; the
    // last entry (the vacuum) is skipped, the operators act right to left, each
    // one contributing (-1)^(occupied modes below its mode) in the blocked
    // order alpha 0..n-1, beta n..2n-1.
    const YamlNode &sup = integrals.at("initial_state_suggestions").at(0).at("state").at("superposition");
    if (sup.kind != YamlNode::kSeq || sup.seq.empty()) fail(sup.line, "superposition must be a non-empty sequence");
    const unsigned n = maxorb;
    uint32_t tab[kBinomRows * kBinomCols];
    fill_binomials(tab);
    int electrons = -1;
    for (const YamlNode &basis : sup.seq) {
        if (basis.kind != YamlNode::kSeq || basis.seq.size() < 2)
            fail(basis.line, "a superposition entry is [coefficient, operators..., vacuum]");
        double coeff = basis.seq[0].as_double();
        if (!std::isfinite(coeff)) fail(basis.line, "non-finite coefficient");
        std::vector<LadderOp> ops;
        for (size_t k = 1; k + 1 < basis.seq.size(); ++k) ops.push_back(parse_ladder(basis.seq[k]));
        uint64_t bv = 0;
        for (size_t k = ops.size(); k-- > 0;) {
            const LadderOp &op = ops[k];
            if (op.orbital >= n)
                fail(basis.line, "initial state names orbital " + std::to_string(op.orbital + 1) +
                                     " beyond the integrals' " + std::to_string(n));
            const unsigned mode = op.up ? op.orbital : op.orbital + n;
            const uint64_t bit = 1ull << mode;
            if (op.raising == ((bv & bit) != 0))
                fail(basis.line, op.raising ? "creation on an occupied spin orbital" : "annihilation on an empty spin orbital");
            bv ^= bit;
            if (popcount64(bv & (bit - 1)) & 1) coeff = -coeff;
        }
        const uint32_t alpha = (uint32_t)(bv & ((1ull << n) - 1));
        const uint32_t beta = (uint32_t)(bv >> n);
        const int na = popcount64(alpha), nb = popcount64(beta);
        if (na != nb)
            fail(basis.line, "determinant has " + std::to_string(na) + " alpha and " + std::to_string(nb) +
                                 " beta electrons; the register holds equal counts");
        if (electrons >= 0 && na != electrons) fail(basis.line, "determinants with different electron counts");
        electrons = na;
        if (na < 1 || (unsigned)na > n || na >= kBinomCols) fail(basis.line, "unsupported electron count");
        const uint64_t C = binom(tab, (int)n, na);
        d.initial.push_back({C * rank_of(alpha, tab) + rank_of(beta, tab), coeff});
    }
    d.num_occupied = (unsigned)electrons;
    return d;
}

void apply_broombridge(const BroombridgeData &d, Hamiltonian &h) {
    for (const auto &t : d.one_body) h.AddOneBodyTerm((int)t.first.first, (int)t.first.second, t.second);
    for (const auto &t : d.two_body)
        h.AddTwoBodyTerm((int)t.first[0], (int)t.first[1], (int)t.first[2], (int)t.first[3], t.second);
    h.SetNumOrbitals(d.num_orbitals);
    h.SetNumOccupied(d.num_occupied);
}

}  // namespace qpe
