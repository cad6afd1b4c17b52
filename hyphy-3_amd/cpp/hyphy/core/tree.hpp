// Phylogenetic trees with HyPhy 3 node numbering.
//
// Mirrors hyphy::core::Tree / TreeNode (include/hyphy/core/tree.hpp:22-892 of
// the reference) for the likelihood path: Newick parsing with node ids in
// creation order (internal node at '(', leaf at ',' or ')'), automatic
// NodeK / LeafK names, {model tags}, [comments], quoted labels, the
// degree-2-root collapse of unroot_if_bifurcating, recursive post-order and
// Newick output.  Reticulate ("Wheatstone bridge") trees are not supported.
#pragma once

#include <algorithm>
#include <cstdint>
#include <iomanip>
#include <sstream>
#include <stack>
#include <stdexcept>
#include <string>
#include <string_view>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "hyphy/core/nexus.hpp"

namespace hyphy::core {

constexpr int32_t INVALID_INDEX = -1;

struct TreeNode {
    int32_t id = INVALID_INDEX;
    std::string name;
    int32_t parent_id = INVALID_INDEX;
    std::vector<int32_t> children;
    double branch_length = 0.0;
    bool is_leaf = false;
    std::string model_tag;
};

class Tree {
public:
    std::vector<TreeNode> nodes;
    int32_t root_id = INVALID_INDEX;
    std::vector<int32_t> post_order;
    std::unordered_map<std::string, int32_t> leaf_name_to_id;

    size_t num_nodes() const { return nodes.size(); }
    size_t num_leaves() const { return leaf_name_to_id.size(); }
    bool has_wheatstone_bridges() const { return false; }

    static Tree from_newick(std::string_view newick_sv, bool keep_rooted = true) {
        std::string newick(newick_sv);
        {
            size_t p = 0;
            while (p < newick.size() && std::isspace((unsigned char)newick[p])) ++p;
            if (nexus::find_ci(newick, "#NEXUS") == p) {
                std::string t = nexus::extract_tree_newick(nexus::strip_comments(newick), 0);
                if (!t.empty()) newick = t;
            }
        }
        while (!newick.empty() && std::strchr("; \t\n\r", newick.back())) newick.pop_back();
        // Drop [comments] and whitespace outside quotes; keep {tags}.
        std::string clean;
        clean.reserve(newick.size());
        {
            bool in_comment = false;
            char q = 0;
            for (size_t k = 0; k < newick.size(); ++k) {
                const char c = newick[k];
                if (q) {
                    clean += c;
                    if (c == q) {
                        if (q == '\'' && k + 1 < newick.size() && newick[k + 1] == '\'')
                            clean += newick[++k];
                        else
                            q = 0;
                    }
                    continue;
                }
                if (in_comment) {
                    if (c == ']') in_comment = false;
                    continue;
                }
                if (c == '[') {
                    in_comment = true;
                    continue;
                }
                if (c == '\'' || c == '"') {
                    q = c;
                    clean += c;
                    continue;
                }
                if (!std::isspace((unsigned char)c)) clean += c;
            }
        }
        Tree tree;
        auto create = [&]() {
            TreeNode n;
            n.id = int32_t(tree.nodes.size());
            tree.nodes.push_back(n);
            return n.id;
        };
        std::stack<int32_t> st;
        std::string token;
        char q = 0;
        for (size_t i = 0; i < clean.size(); ++i) {
            const char c = clean[i];
            if (q) {
                token += c;
                if (c == q) {
                    if (q == '\'' && i + 1 < clean.size() && clean[i + 1] == '\'')
                        token += clean[++i];
                    else
                        q = 0;
                }
                continue;
            }
            if (c == '\'' || c == '"') {
                q = c;
                token += c;
                continue;
            }
            if (c == ';') break;
            if (c == '(') {
                const int32_t id = create();
                if (!st.empty()) {
                    tree.nodes[st.top()].children.push_back(id);
                    tree.nodes[id].parent_id = st.top();
                }
                st.push(id);
            } else if (c == ',' || c == ')') {
                if (!token.empty()) {
                    const int32_t id = create();
                    tree.nodes[id].is_leaf = true;
                    apply_token(tree.nodes[id], token);
                    token.clear();
                    if (!st.empty()) {
                        tree.nodes[st.top()].children.push_back(id);
                        tree.nodes[id].parent_id = st.top();
                    }
                }
                if (c == ')') {
                    if (st.empty()) throw std::runtime_error("Invalid Newick tree: unbalanced parentheses");
                    const int32_t cur = st.top();
                    st.pop();
                    size_t j = i + 1;
                    std::string it;
                    char iq = 0;
                    while (j < clean.size()) {
                        const char cj = clean[j];
                        if (iq) {
                            it += cj;
                            if (cj == iq) {
                                if (iq == '\'' && j + 1 < clean.size() && clean[j + 1] == '\'')
                                    it += clean[++j];
                                else
                                    iq = 0;
                            }
                            ++j;
                            continue;
                        }
                        if (cj == '\'' || cj == '"') {
                            iq = cj;
                            it += cj;
                            ++j;
                            continue;
                        }
                        if (cj == ',' || cj == ')' || cj == '(' || cj == ';') break;
                        it += cj;
                        ++j;
                    }
                    if (iq) throw std::runtime_error("Invalid Newick tree: unterminated quoted label");
                    if (!it.empty()) {
                        apply_token(tree.nodes[cur], it);
                        i = j - 1;
                    }
                }
            } else {
                token += c;
            }
        }
        if (q) throw std::runtime_error("Invalid Newick tree: unterminated quoted label");
        if (!st.empty()) throw std::runtime_error("Invalid Newick tree: unbalanced parentheses");
        if (tree.nodes.empty()) throw std::runtime_error("Failed to parse Newick tree: empty tree");
        std::unordered_set<std::string> used;
        for (const auto &n : tree.nodes)
            if (n.is_leaf && !n.name.empty()) used.insert(n.name);
        int counter = 1;
        for (auto &n : tree.nodes) {
            if (n.parent_id == INVALID_INDEX) tree.root_id = n.id;
            if (!n.is_leaf) {
                if (n.name.empty() || used.count(n.name)) {
                    std::string cand;
                    do cand = "Node" + std::to_string(counter++);
                    while (used.count(cand));
                    n.name = cand;
                }
                used.insert(n.name);
            } else if (n.name.empty()) {
                std::string cand;
                do cand = "Leaf" + std::to_string(counter++);
                while (used.count(cand));
                n.name = cand;
                used.insert(n.name);
            }
        }
        if (!keep_rooted) tree.unroot_if_bifurcating();
        tree.rebuild_leaf_map();
        tree.compute_post_order();
        return tree;
    }

    // Collapse a degree-2 root into its internal child (tree.hpp:591-690).
    void unroot_if_bifurcating() {
        if (root_id == INVALID_INDEX || nodes.size() <= 3) return;
        if (nodes[root_id].children.size() != 2) return;
        const int32_t c0 = nodes[root_id].children[0], c1 = nodes[root_id].children[1];
        int32_t absorb, other;
        if (!nodes[c0].is_leaf && !nodes[c1].is_leaf) {
            if (nodes[c1].model_tag.empty() && !nodes[c0].model_tag.empty()) {
                absorb = c1;
                other = c0;
            } else {
                absorb = c0;
                other = c1;
            }
        } else if (!nodes[c0].is_leaf) {
            absorb = c0;
            other = c1;
        } else if (!nodes[c1].is_leaf) {
            absorb = c1;
            other = c0;
        } else {
            return;
        }
        nodes[other].branch_length += nodes[absorb].branch_length;
        if (nodes[other].model_tag.empty() && !nodes[absorb].model_tag.empty()) nodes[other].model_tag = nodes[absorb].model_tag;
        std::vector<int32_t> rc;
        if (absorb == c0) {
            rc = nodes[absorb].children;
            rc.push_back(other);
        } else {
            rc.push_back(other);
            rc.insert(rc.end(), nodes[absorb].children.begin(), nodes[absorb].children.end());
        }
        nodes[root_id].children = rc;
        const int32_t old_n = int32_t(nodes.size());
        std::vector<int32_t> map(old_n, INVALID_INDEX);
        std::vector<TreeNode> nn;
        nn.reserve(old_n - 1);
        for (int32_t i = 0; i < old_n; ++i) {
            if (i == absorb) continue;
            map[i] = int32_t(nn.size());
            TreeNode t = std::move(nodes[i]);
            t.id = map[i];
            nn.push_back(std::move(t));
        }
        root_id = map[root_id];
        for (auto &t : nn) {
            if (t.id == root_id) {
                t.parent_id = INVALID_INDEX;
                t.branch_length = 0.0;
            }
            for (auto &c : t.children) {
                c = map[c];
                if (c != INVALID_INDEX) nn[c].parent_id = t.id;
            }
        }
        nodes = std::move(nn);
        rebuild_leaf_map();
        compute_post_order();
    }

    void rebuild_leaf_map() {
        leaf_name_to_id.clear();
        for (const auto &n : nodes)
            if (n.is_leaf) leaf_name_to_id[n.name] = n.id;
    }

    void compute_post_order() {
        post_order.clear();
        if (root_id == INVALID_INDEX) return;
        std::vector<std::pair<int32_t, size_t>> st{{root_id, 0}};
        while (!st.empty()) {
            auto &[n, k] = st.back();
            if (k < nodes[n].children.size()) {
                const int32_t c = nodes[n].children[k++];
                st.push_back({c, 0});
            } else {
                post_order.push_back(n);
                st.pop_back();
            }
        }
    }

    std::string to_newick(bool include_branch_lengths = true) const {
        if (root_id == INVALID_INDEX) return "";
        std::ostringstream os;
        write(os, root_id, include_branch_lengths);
        os << ";";
        return os.str();
    }

private:
    void write(std::ostringstream &os, int32_t id, bool bl) const {
        const auto &n = nodes[id];
        if (!n.children.empty()) {
            os << "(";
            for (size_t i = 0; i < n.children.size(); ++i) {
                if (i) os << ",";
                write(os, n.children[i], bl);
            }
            os << ")";
        }
        os << n.name;
        if (!n.model_tag.empty()) os << "{" << n.model_tag << "}";
        if (bl && id != root_id) os << ":" << std::setprecision(17) << n.branch_length;
    }

    static double stod_prefix(const std::string &s) {
        try {
            return std::stod(s);
        } catch (...) {
            return 0.0;
        }
    }

    static void apply_token(TreeNode &node, std::string str) {
        if (str.empty()) return;
        std::string tag;
        size_t ob = std::string::npos, cb = std::string::npos;
        {
            char q = 0;
            for (size_t k = 0; k < str.size(); ++k) {
                const char ch = str[k];
                if (q) {
                    if (ch == q) {
                        if (q == '\'' && k + 1 < str.size() && str[k + 1] == '\'')
                            ++k;
                        else
                            q = 0;
                    }
                    continue;
                }
                if (ch == '\'' || ch == '"') {
                    q = ch;
                    continue;
                }
                if (ch == '{' && ob == std::string::npos) {
                    ob = k;
                } else if (ch == '}' && ob != std::string::npos && cb == std::string::npos) {
                    cb = k;
                    break;
                }
            }
        }
        if (ob != std::string::npos && cb != std::string::npos && cb > ob) {
            tag = str.substr(ob + 1, cb - ob - 1);
            str.erase(ob, cb - ob + 1);
        }
        size_t colon = std::string::npos;
        {
            char q = 0;
            for (size_t k = 0; k < str.size(); ++k) {
                const char ch = str[k];
                if (q) {
                    if (ch == q) {
                        if (q == '\'' && k + 1 < str.size() && str[k + 1] == '\'')
                            ++k;
                        else
                            q = 0;
                    }
                    continue;
                }
                if (ch == '\'' || ch == '"') {
                    q = ch;
                    continue;
                }
                if (ch == ':') {
                    colon = k;
                    break;
                }
            }
        }
        std::string name = colon != std::string::npos ? str.substr(0, colon) : str;
        const double len = colon != std::string::npos ? stod_prefix(str.substr(colon + 1)) : 0.0;
        if (name.size() >= 2 && ((name.front() == '\'' && name.back() == '\'') || (name.front() == '"' && name.back() == '"'))) {
            const char qq = name.front();
            std::string inner = name.substr(1, name.size() - 2), un;
            if (qq == '\'') {
                for (size_t k = 0; k < inner.size(); ++k) {
                    if (inner[k] == '\'' && k + 1 < inner.size() && inner[k + 1] == '\'') {
                        un += '\'';
                        ++k;
                    } else {
                        un += inner[k];
                    }
                }
                name = un;
            } else {
                name = inner;
            }
        }
        node.name = name;
        node.branch_length = len;
        node.model_tag = tag;
    }
};

}  // namespace hyphy::core
