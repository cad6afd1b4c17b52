// Minimal NEXUS helpers for alignment and tree input: comment stripping,
// case-insensitive keyword search, quoted labels and the first TREE of a
// TREES block with TRANSLATE applied (semantics of
// include/hyphy/core/nexus_utils.hpp:13-455 of the reference).
#pragma once

#include <cctype>
#include <cstring>
#include <string>
#include <unordered_map>

namespace hyphy::core::nexus {

// Remove [comments] (nested) outside quotes, keeping line breaks.
inline std::string strip_comments(const std::string &raw) {
    std::string out;
    out.reserve(raw.size());
    int depth = 0;
    char q = 0;
    for (size_t i = 0; i < raw.size(); ++i) {
        const char c = raw[i];
        if (depth > 0) {
            if (c == '[')
                ++depth;
            else if (c == ']')
                --depth;
            else if (c == '\n' || c == '\r')
                out += c;
            continue;
        }
        if (q) {
            out += c;
            if (c == q) {
                if (c == '\'' && i + 1 < raw.size() && raw[i + 1] == '\'')
                    out += raw[++i];
                else
                    q = 0;
            } else if (c == '\n' || c == '\r') {
                q = 0;
            }
            continue;
        }
        if (c == '[') {
            ++depth;
        } else {
            if (c == '\'' || c == '"') q = c;
            out += c;
        }
    }
    return out;
}

inline size_t find_ci(const std::string &hay, const std::string &needle, size_t from = 0) {
    if (needle.empty() || hay.size() < needle.size()) return std::string::npos;
    for (size_t i = from; i + needle.size() <= hay.size(); ++i) {
        size_t k = 0;
        while (k < needle.size() && std::toupper((unsigned char)hay[i + k]) == std::toupper((unsigned char)needle[k])) ++k;
        if (k == needle.size()) return i;
    }
    return std::string::npos;
}

// Keyword as a whole word (delimited by whitespace, ';' or text bounds).
inline size_t find_word_ci(const std::string &t, const std::string &kw, size_t from, size_t to) {
    for (size_t p = find_ci(t, kw, from); p != std::string::npos && p < to; p = find_ci(t, kw, p + 1)) {
        const bool lb = p == 0 || std::isspace((unsigned char)t[p - 1]) || t[p - 1] == ';';
        const size_t e = p + kw.size();
        const bool rb = e >= t.size() || std::isspace((unsigned char)t[e]) || t[e] == ';';
        if (lb && rb) return p;
    }
    return std::string::npos;
}

// Read a (possibly quoted) label at position i; advances i past it.
inline std::string read_label(const std::string &s, size_t &i) {
    std::string out;
    if (i < s.size() && (s[i] == '\'' || s[i] == '"')) {
        const char q = s[i++];
        while (i < s.size()) {
            if (s[i] == q) {
                if (q == '\'' && i + 1 < s.size() && s[i + 1] == '\'') {
                    out += '\'';
                    i += 2;
                    continue;
                }
                ++i;
                break;
            }
            out += s[i++];
        }
        return out;
    }
    while (i < s.size() && !std::isspace((unsigned char)s[i]) && s[i] != ',' && s[i] != ';') out += s[i++];
    return out;
}

inline std::string quote_label_if_needed(const std::string &lbl) {
    bool plain = !lbl.empty();
    for (char c : lbl)
        if (std::isspace((unsigned char)c) || std::strchr("(),:;[]'\"{}", c)) plain = false;
    if (plain) return lbl;
    std::string o = "'";
    for (char c : lbl) {
        if (c == '\'') o += '\'';
        o += c;
    }
    return o + "'";
}

// First TREE/UTREE statement after `from` (a BEGIN TREES block is preferred),
// with TRANSLATE keys replaced by their labels.
inline std::string extract_tree_newick(const std::string &t, size_t from = 0) {
    size_t scope_start = 0, scope_end = t.size();
    size_t bt = std::string::npos;
    for (size_t p = find_ci(t, "BEGIN", 0); p != std::string::npos; p = find_ci(t, "BEGIN", p + 5)) {
        size_t q = p + 5;
        while (q < t.size() && std::isspace((unsigned char)t[q])) ++q;
        if (find_ci(t, "TREES", q) == q) {
            bt = q + 5;
            break;
        }
    }
    if (bt != std::string::npos) {
        const size_t semi = t.find(';', bt);
        scope_start = semi == std::string::npos ? bt : semi + 1;
        const size_t eb = find_word_ci(t, "ENDBLOCK", scope_start, t.size());
        const size_t ee = find_word_ci(t, "END", scope_start, t.size());
        scope_end = (eb != std::string::npos && (ee == std::string::npos || eb < ee)) ? eb
                    : (ee != std::string::npos ? ee : t.size());
    } else {
        scope_start = from;
    }
    std::unordered_map<std::string, std::string> tr;
    const size_t tp = find_word_ci(t, "TRANSLATE", scope_start, scope_end);
    if (tp != std::string::npos) {
        size_t i = tp + 9;
        const size_t semi = t.find(';', i);
        const size_t lim = semi == std::string::npos ? scope_end : semi;
        const std::string body = t.substr(i, lim - i);
        size_t k = 0;
        while (k < body.size()) {
            while (k < body.size() && (std::isspace((unsigned char)body[k]) || body[k] == ',')) ++k;
            if (k >= body.size()) break;
            std::string key = read_label(body, k);
            while (k < body.size() && std::isspace((unsigned char)body[k])) ++k;
            if (k >= body.size()) break;
            std::string val = read_label(body, k);
            if (!key.empty()) tr[key] = val;
        }
    }
    size_t tree = find_word_ci(t, "TREE", scope_start, scope_end);
    const size_t utree = find_word_ci(t, "UTREE", scope_start, scope_end);
    if (utree != std::string::npos && (tree == std::string::npos || utree < tree)) tree = utree;
    if (tree == std::string::npos) return "";
    const size_t eq = t.find('=', tree);
    if (eq == std::string::npos || eq >= scope_end) return "";
    const size_t lp = t.find('(', eq);
    if (lp == std::string::npos || lp >= scope_end) return "";
    int depth = 0;
    size_t r = lp;
    char q = 0;
    for (; r < scope_end; ++r) {
        const char c = t[r];
        if (q) {
            if (c == q) {
                if (q == '\'' && r + 1 < scope_end && t[r + 1] == '\'')
                    ++r;
                else
                    q = 0;
            }
            continue;
        }
        if (c == '\'' || c == '"')
            q = c;
        else if (c == '(')
            ++depth;
        else if (c == ')' && --depth == 0)
            break;
    }
    if (r >= scope_end) return "";
    size_t semi = t.find(';', r);
    std::string nwk = (semi != std::string::npos && semi < scope_end && semi - r <= 64) ? t.substr(lp, semi - lp + 1)
                                                                                          : t.substr(lp, r - lp + 1) + ";";
    if (tr.empty()) return nwk;
    // Replace leaf tokens (those after '(' or ',') found in the TRANSLATE map.
    std::string out;
    bool expect = false;
    for (size_t i = 0; i < nwk.size();) {
        const char c = nwk[i];
        if (expect && !std::isspace((unsigned char)c) && !std::strchr("(),:;{", c)) {
            size_t j = i;
            std::string tok;
            if (c == '\'' || c == '"') {
                tok = read_label(nwk, j);
            } else {
                while (j < nwk.size() && !std::isspace((unsigned char)nwk[j]) && !std::strchr("():,;{[", nwk[j])) ++j;
                tok = nwk.substr(i, j - i);
            }
            auto it = tr.find(tok);
            out += it != tr.end() ? quote_label_if_needed(it->second) : nwk.substr(i, j - i);
            i = j;
            expect = false;
            continue;
        }
        out += c;
        if (c == '(' || c == ',')
            expect = true;
        else if (!std::isspace((unsigned char)c))
            expect = false;
        ++i;
    }
    return out;
}

}  // namespace hyphy::core::nexus
