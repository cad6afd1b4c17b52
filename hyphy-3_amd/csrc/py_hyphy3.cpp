// _hyphy3: Python module with the reference binding's names for the
// likelihood path (src/python/bindings.cpp:40-119,725-732 of the reference),
// built with pybind11 over the C++ data layer and the MI355X facade.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "hyphy/core/alignment.hpp"
#include "hyphy/core/likelihood.hpp"
#include "hyphy/core/tree.hpp"

namespace py = pybind11;
using namespace hyphy::core;

namespace {

// [node][state][pattern] array of a workspace's per-node S x B matrices
py::array_t<double> node_stack(const std::vector<Matrix> &m) {
    const py::ssize_t N = py::ssize_t(m.size()), S = N ? m[0].rows() : 0, B = N ? m[0].cols() : 0;
    py::array_t<double> a({N, S, B});
    auto w = a.mutable_unchecked<3>();
    for (py::ssize_t n = 0; n < N; ++n)
        for (py::ssize_t i = 0; i < S; ++i)
            for (py::ssize_t b = 0; b < B; ++b) w(n, i, b) = m[n](i, b);
    return a;
}

template <typename T>
py::array_t<T> dense(const DenseMatrix<T> &m) {
    py::array_t<T> a({py::ssize_t(m.rows()), py::ssize_t(m.cols())});
    auto w = a.template mutable_unchecked<2>();
    for (long i = 0; i < m.rows(); ++i)
        for (long j = 0; j < m.cols(); ++j) w(i, j) = m(i, j);
    return a;
}

std::vector<std::vector<double>> p_list(const py::array_t<double, py::array::c_style | py::array::forcecast> &P) {
    if (P.ndim() != 3 || P.shape(1) != P.shape(2)) throw std::invalid_argument("P_branches must be num_nodes x S x S");
    std::vector<std::vector<double>> out(P.shape(0));
    const size_t SS = size_t(P.shape(1)) * P.shape(2);
    for (py::ssize_t n = 0; n < P.shape(0); ++n) out[n].assign(P.data() + n * SS, P.data() + (n + 1) * SS);
    return out;
}

py::dict problem_arrays(const Tree &tree, const Alignment &aln) {
    const auto l2t = LikelihoodEngine::build_leaf_to_taxon_map(tree, aln);
    const size_t N = tree.num_nodes();
    py::array_t<int32_t> parent(N), cp(N + 1), lt(N);
    std::vector<int32_t> ci;
    auto pp = parent.mutable_unchecked<1>();
    auto cpp = cp.mutable_unchecked<1>();
    auto ltp = lt.mutable_unchecked<1>();
    cpp(0) = 0;
    for (size_t n = 0; n < N; ++n) {
        pp(n) = tree.nodes[n].parent_id;
        ltp(n) = (tree.nodes[n].is_leaf && l2t[n] != size_t(-1)) ? int32_t(l2t[n]) : -1;
        for (auto c : tree.nodes[n].children) ci.push_back(c);
        cpp(n + 1) = int32_t(ci.size());
    }
    std::vector<int8_t> st;
    std::vector<uint64_t> mk;
    std::vector<double> w;
    aln.pattern_arrays(st, mk, w);
    const py::ssize_t P = py::ssize_t(aln.patterns.size()), T = py::ssize_t(aln.num_taxa);
    py::array_t<int8_t> states({P, T});
    py::array_t<uint64_t> masks({P, T});
    py::array_t<double> weights(P);
    std::copy(st.begin(), st.end(), states.mutable_data());
    std::copy(mk.begin(), mk.end(), masks.mutable_data());
    std::copy(w.begin(), w.end(), weights.mutable_data());
    py::array_t<int32_t> child_idx(py::ssize_t(ci.size()));
    std::copy(ci.begin(), ci.end(), child_idx.mutable_data());
    py::dict d;
    d["parent"] = parent;
    d["child_ptr"] = cp;
    d["child_idx"] = child_idx;
    d["leaf_taxon"] = lt;
    d["states"] = states;
    d["masks"] = masks;
    d["weights"] = weights;
    // a nucleotide view (no codons) is a 4-state problem with the nucleotide frequencies
    const bool nuc = aln.num_codons == 0;
    d["S"] = nuc ? 4 : (aln.code ? aln.code->num_sense_codons : 61);
    d["code_id"] = aln.code ? aln.code->id : 0;
    py::array_t<double> pos({3, 4});
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 4; ++b) pos.mutable_at(a, b) = aln.pos_nuc_frequencies[a][b];
    d["pos_freqs"] = pos;
    if (nuc)
        d["pi"] = py::array_t<double>(4, aln.nuc_frequencies.data());
    else
        d["pi"] = py::array_t<double>(py::ssize_t(aln.codon_frequencies_f3x4.size()), aln.codon_frequencies_f3x4.data());
    return d;
}

}  // namespace

PYBIND11_MODULE(_hyphy3, m) {
    m.doc() = "HyPhy 3 codon likelihood core on MI355X (MG94xREV pruning + inside-outside adjoint)";

    py::class_<GeneticCode, std::shared_ptr<GeneticCode>>(m, "GeneticCode")
        .def_static("universal", []() { return std::const_pointer_cast<GeneticCode>(GeneticCode::universal()); })
        .def_static("from_name", [](const std::string &n) { return std::const_pointer_cast<GeneticCode>(GeneticCode::from_name(n)); },
                    py::arg("name"))
        .def_readonly("name", &GeneticCode::name)
        .def_readonly("id", &GeneticCode::id)
        .def_readonly("num_sense_codons", &GeneticCode::num_sense_codons)
        .def_readonly("sense_codons", &GeneticCode::sense_codons)
        .def_readonly("stop_codons", &GeneticCode::stop_codons)
        .def("is_synonymous", &GeneticCode::is_synonymous)
        .def_static("from_id", [](int id) { return std::const_pointer_cast<GeneticCode>(GeneticCode::from_id(id)); },
                    py::arg("id"))
        // amino-acid letter of every codon in ACGT order ('*' = stop)
        .def_property_readonly("codon_aa64", [](const GeneticCode &g) { return std::string(g.codon_aa64, 64); })
        // diff_matrix entry of sense pair (u, v): (num_diffs, synonymous,
        // diff_pos[3], diff_from[3], diff_to[3])
        .def("diff", [](const GeneticCode &g, int u, int v) {
            if (u < 0 || v < 0 || u >= g.num_sense_codons || v >= g.num_sense_codons)
                throw py::index_error("sense codon index out of range");
            const auto &d = g.diff_matrix[u][v];
            auto tri = [](const int8_t *x) { return py::make_tuple(int(x[0]), int(x[1]), int(x[2])); };
            return py::make_tuple(int(d.num_diffs), bool(d.synonymous), tri(d.diff_pos), tri(d.diff_from),
                                  tri(d.diff_to));
        }, py::arg("u"), py::arg("v"));

    py::class_<TreeNode>(m, "TreeNode")
        .def_readonly("id", &TreeNode::id)
        .def_readonly("name", &TreeNode::name)
        .def_readonly("parent_id", &TreeNode::parent_id)
        .def_readonly("children", &TreeNode::children)
        .def_readwrite("branch_length", &TreeNode::branch_length)
        .def_readonly("is_leaf", &TreeNode::is_leaf)
        .def_readonly("model_tag", &TreeNode::model_tag);

    py::class_<Tree>(m, "Tree")
        .def(py::init<>())
        .def_static("from_newick", [](const std::string &s, bool keep_rooted) { return Tree::from_newick(s, keep_rooted); },
                    py::arg("newick"), py::arg("keep_rooted") = true)
        .def_static("from_newick_file",
                    [](const std::string &path, bool keep_rooted) {
                        return Tree::from_newick(Alignment::read_file(path), keep_rooted);
                    },
                    py::arg("filepath"), py::arg("keep_rooted") = true)
        .def("num_nodes", &Tree::num_nodes)
        .def("num_leaves", &Tree::num_leaves)
        .def_readonly("root_id", &Tree::root_id)
        .def_readonly("post_order", &Tree::post_order)
        .def_readonly("nodes", &Tree::nodes)
        .def_readonly("leaf_name_to_id", &Tree::leaf_name_to_id)
        .def("to_newick", &Tree::to_newick, py::arg("include_branch_lengths") = true)
        .def("get_branch_lengths",
             [](const Tree &t) {
                 std::vector<double> bl(t.num_nodes());
                 for (size_t i = 0; i < t.nodes.size(); ++i) bl[i] = t.nodes[i].branch_length;
                 return bl;
             })
        .def("set_branch_lengths",
             [](Tree &t, const std::vector<double> &bl) {
                 for (size_t i = 0; i < t.nodes.size() && i < bl.size(); ++i) t.nodes[i].branch_length = bl[i];
             })
        .def("get_branch_length", [](const Tree &t, int32_t id) { return t.nodes.at(id).branch_length; }, py::arg("id"))
        .def("set_branch_length", [](Tree &t, int32_t id, double v) { t.nodes.at(id).branch_length = v; }, py::arg("id"),
             py::arg("length"));

    py::class_<Alignment>(m, "Alignment")
        .def(py::init<>())
        .def_static("load", [](const std::string &p) { return Alignment::load(p); }, py::arg("filepath"))
        .def_static("from_fasta", [](const std::string &p) { return Alignment::from_fasta(p); }, py::arg("filepath"))
        .def_static("from_nexus", [](const std::string &p) { return Alignment::from_nexus(p); }, py::arg("filepath"))
        .def_static("from_text",
                    [](const std::string &text, const std::string &code) {
                        return Alignment::load_text(text, GeneticCode::from_name(code));
                    },
                    py::arg("text"), py::arg("genetic_code") = "Universal")
        .def_readonly("num_taxa", &Alignment::num_taxa)
        .def_property_readonly("num_sequences", [](const Alignment &a) { return a.num_taxa; })
        .def_readonly("num_codons", &Alignment::num_codons)
        .def_readonly("num_nucleotides", &Alignment::num_nucleotides)
        .def_property_readonly("num_sites", [](const Alignment &a) { return a.num_codons; })
        .def_property_readonly("num_patterns", [](const Alignment &a) { return a.patterns.size(); })
        .def_readonly("embedded_tree_newick", &Alignment::embedded_tree_newick)
        .def_readonly("taxon_names", &Alignment::taxon_names)
        .def_readonly("site_to_pattern", &Alignment::site_to_pattern)
        .def_readonly("nuc_frequencies", &Alignment::nuc_frequencies)
        .def("to_nucleotide_alignment", &Alignment::to_nucleotide_alignment,
             "Nucleotide view for the GTR stage (one site per nucleotide column)")
        .def_readonly("pos_nuc_frequencies", &Alignment::pos_nuc_frequencies)
        .def_readonly("codon_frequencies_f1x4", &Alignment::codon_frequencies_f1x4)
        .def_readonly("codon_frequencies_f3x4", &Alignment::codon_frequencies_f3x4)
        .def_property_readonly("pattern_weights",
                               [](const Alignment &a) {
                                   std::vector<size_t> w;
                                   for (const auto &p : a.patterns) w.push_back(p.weight);
                                   return w;
                               })
        .def_property_readonly("pattern_sites",
                               [](const Alignment &a) {
                                   std::vector<std::vector<size_t>> s;
                                   for (const auto &p : a.patterns) s.push_back(p.sites);
                                   return s;
                               });

    py::class_<MG94Parameters>(m, "MG94Parameters")
        .def(py::init<>())
        .def_readwrite("alpha", &MG94Parameters::alpha)
        .def_readwrite("beta", &MG94Parameters::beta)
        .def_readwrite("theta_AC", &MG94Parameters::theta_AC)
        .def_readwrite("theta_AT", &MG94Parameters::theta_AT)
        .def_readwrite("theta_CG", &MG94Parameters::theta_CG)
        .def_readwrite("theta_CT", &MG94Parameters::theta_CT)
        .def_readwrite("theta_GT", &MG94Parameters::theta_GT)
        .def_readwrite("delta", &MG94Parameters::delta)
        .def_readwrite("psi", &MG94Parameters::psi);

    m.def("set_precision",
          [](const std::string &p) {
              if (p == "fp32")
                  LikelihoodEngine::set_precision(LikelihoodEngine::FP32);
              else if (p == "fp64")
                  LikelihoodEngine::set_precision(LikelihoodEngine::FP64);
              else
                  throw std::invalid_argument("precision must be 'fp32' or 'fp64'");
          },
          py::arg("precision"));
    m.def("get_precision", []() { return LikelihoodEngine::get_precision() == LikelihoodEngine::FP32 ? "fp32" : "fp64"; });

    m.def("compute_branch_length_gradients",
          [](const Tree &tree, const Alignment &aln, const MG94Parameters &params, const std::vector<double> &bl) {
              py::gil_scoped_release nogil;
              return LikelihoodEngine::compute_branch_length_gradients(tree, aln, params, bl);
          },
          py::arg("tree"), py::arg("aln"), py::arg("params"), py::arg("branch_lengths") = std::vector<double>{},
          "Compute log-likelihood and exact analytical branch gradients d ln L / d t_b via Inside-Outside");

    m.def("compute_mg94_log_likelihood",
          [](const Tree &tree, const Alignment &aln, const MG94Parameters &params, double branch_scale) {
              MG94Matrix mg;  // the evaluation builds its own spectra on the device
              mg.update(params, aln.pos_nuc_frequencies, aln.codon_frequencies_f3x4, *aln.code, false);
              py::gil_scoped_release nogil;
              return LikelihoodEngine::compute_mg94_log_likelihood(tree, aln, mg, nullptr, branch_scale);
          },
          py::arg("tree"), py::arg("aln"), py::arg("params"), py::arg("branch_scale") = 1.0);

    m.def("mg94_matrix",
          [](const Alignment &aln, const MG94Parameters &params) {
              MG94Matrix mg;
              {
                  py::gil_scoped_release nogil;
                  mg.update(params, aln.pos_nuc_frequencies, aln.codon_frequencies_f3x4, *aln.code);
              }
              py::dict d;
              d["Q"] = dense(mg.Q);
              d["eigenvalues"] = mg.eigenvalues.v;
              d["eigenvectors"] = dense(mg.eigenvectors);
              d["U_pi"] = dense(mg.U_pi);
              d["V_pi"] = dense(mg.V_pi);
              d["scale_factor"] = mg.scale_factor;
              return d;
          },
          py::arg("aln"), py::arg("params"),
          "MG94Matrix::update on the alignment's frequencies: Q, spectrum (GPU), U_pi, V_pi, scale_factor");

    m.def("mg94_transition_matrix",
          [](const Alignment &aln, const MG94Parameters &params, double t) {
              MG94Matrix mg;
              Matrix P;
              {
                  py::gil_scoped_release nogil;
                  mg.update(params, aln.pos_nuc_frequencies, aln.codon_frequencies_f3x4, *aln.code);
                  P = mg.transition_matrix(t);
              }
              return dense(P);
          },
          py::arg("aln"), py::arg("params"), py::arg("t"), "MG94Matrix::transition_matrix(t)");

    m.def("compute_joint_frechet_gradients",
          [](const Tree &tree, const Alignment &aln, const MG94Parameters &params, const std::vector<double> &bl,
             py::object is_test, double omega_bg) {
              std::vector<bool> it;
              const std::vector<bool> *itp = nullptr;
              if (!is_test.is_none()) {
                  it = is_test.cast<std::vector<bool>>();
                  itp = &it;
              }
              JointMG94Gradients r;
              {
                  py::gil_scoped_release nogil;
                  r = LikelihoodEngine::compute_joint_frechet_gradients(tree, aln, params, bl, nullptr, itp, omega_bg);
              }
              py::dict d;
              d["log_likelihood"] = r.log_likelihood;
              d["grad_branches"] = r.grad_branches;
              d["grad_omega"] = r.grad_omega;
              d["grad_omega_bg"] = r.grad_omega_bg;
              d["grad_theta"] = std::vector<double>{r.grad_theta_AC, r.grad_theta_AT, r.grad_theta_CG, r.grad_theta_CT,
                                                    r.grad_theta_GT};
              return d;
          },
          py::arg("tree"), py::arg("aln"), py::arg("params"), py::arg("branch_lengths") = std::vector<double>{},
          py::arg("is_test_branch") = py::none(), py::arg("omega_bg") = -1.0);

    m.def("compute_mixture_gradients",
          [](const Tree &tree, const Alignment &aln, const MG94Parameters &params, const std::vector<double> &omegas,
             const std::vector<double> &weights, const std::vector<double> &bl, int time_convention, bool want_grad) {
              MixtureGradients r;
              {
                  py::gil_scoped_release nogil;
                  r = LikelihoodEngine::compute_mixture_gradients(tree, aln, params, omegas, weights, bl, time_convention,
                                                                  want_grad);
              }
              py::dict d;
              d["log_likelihood"] = r.log_likelihood;
              d["site_log_likelihoods"] = r.site_log_likelihoods;
              d["grad_branches"] = r.grad_branches;
              d["grad_omegas"] = r.grad_omegas;
              d["grad_weights"] = r.grad_weights;
              d["grad_theta"] = std::vector<double>(r.grad_theta.begin(), r.grad_theta.end());
              return d;
          },
          py::arg("tree"), py::arg("aln"), py::arg("params"), py::arg("omegas"), py::arg("weights"),
          py::arg("branch_lengths") = std::vector<double>{}, py::arg("time_convention") = 0, py::arg("want_grad") = true);

    m.def("compute_inside_batch_in_place",
          [](const Tree &tree, const Alignment &aln, size_t p_start, size_t p_end,
             const py::array_t<double, py::array::c_style | py::array::forcecast> &P, const std::vector<double> &pi) {
              const auto l2t = LikelihoodEngine::build_leaf_to_taxon_map(tree, aln);
              const auto Pl = p_list(P);
              LikelihoodEngine::BatchedInsideWorkspace ws;
              {
                  py::gil_scoped_release nogil;
                  LikelihoodEngine::compute_inside_batch_in_place(tree, aln, p_start, p_end, l2t, Pl, pi, int(pi.size()), ws);
              }
              py::dict d;
              d["D"] = node_stack(ws.D);
              d["node_log_scale"] = dense(ws.node_log_scale);
              d["root_L"] = ws.root_L.v;
              d["root_log_L"] = ws.root_log_L.v;
              return d;
          },
          py::arg("tree"), py::arg("aln"), py::arg("p_start"), py::arg("p_end"), py::arg("P_branches"), py::arg("pi"),
          "Workspace of compute_inside_batch_in_place (arrays [node][state][pattern])");

    m.def("compute_inside_outside_batch_in_place",
          [](const Tree &tree, const Alignment &aln, size_t p_start, size_t p_end,
             const py::array_t<double, py::array::c_style | py::array::forcecast> &P, const std::vector<double> &pi) {
              const auto l2t = LikelihoodEngine::build_leaf_to_taxon_map(tree, aln);
              const auto Pl = p_list(P);
              LikelihoodEngine::BatchedInsideOutsideWorkspace ws;
              {
                  py::gil_scoped_release nogil;
                  LikelihoodEngine::compute_inside_outside_batch_in_place(tree, aln, p_start, p_end, l2t, Pl, pi,
                                                                          int(pi.size()), ws);
              }
              py::dict d;
              d["D"] = node_stack(ws.D);
              d["V"] = node_stack(ws.V);
              d["P_D"] = node_stack(ws.P_D);
              d["U"] = node_stack(ws.U);
              d["boost_counts"] = dense(ws.boost_counts);
              d["node_log_scale"] = dense(ws.node_log_scale);
              d["likelihood"] = ws.likelihood.v;
              d["log_likelihood"] = ws.log_likelihood.v;
              return d;
          },
          py::arg("tree"), py::arg("aln"), py::arg("p_start"), py::arg("p_end"), py::arg("P_branches"), py::arg("pi"),
          "Workspace of compute_inside_outside_batch_in_place (arrays [node][state][pattern])");

    m.def("compute_gtr_log_likelihood",
          [](const Tree &tree, const Alignment &aln, const std::vector<double> &theta, const std::vector<double> &freqs) {
              if (theta.size() != 5 || freqs.size() != 4)
                  throw std::invalid_argument("theta: 5 exchangeabilities (AC, AT, CG, CT, GT); freqs: 4 frequencies");
              GTRParameters p;
              p.theta_AC = theta[0];
              p.theta_AT = theta[1];
              p.theta_CG = theta[2];
              p.theta_CT = theta[3];
              p.theta_GT = theta[4];
              GTRMatrix g;
              g.update(p, {freqs[0], freqs[1], freqs[2], freqs[3]});
              const Alignment nuc = aln.num_codons == 0 ? aln : aln.to_nucleotide_alignment();
              return LikelihoodEngine::compute_gtr_log_likelihood(tree, nuc, g);
          },
          py::arg("tree"), py::arg("aln"), py::arg("theta"), py::arg("freqs"),
          "GTR nucleotide lnL (compute_gtr_log_likelihood) of the alignment's nucleotide view");
    m.def("problem_arrays", &problem_arrays, py::arg("tree"), py::arg("aln"),
          "Flat tree / pattern arrays in the layout of the C ABI's hy_set_problem and hy_set_model");
}
