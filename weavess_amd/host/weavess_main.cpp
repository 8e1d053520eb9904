// weavess_main.cpp -- KGraph, EFANNA, NSG, SSG, Vamana and DPG drivers of
// the pipeline (test/main.cpp:6-34, :66-95, :97-126, :128-157, :159-188,
// :190-217 and :480-546 of the reference, with the HIP stages).
//
//   weavess_main kgraph|efanna|nsg|ssg|vamana|dpg <prefix> build|search|info [L_search]
//
// <prefix>_base.fvecs, <prefix>_query.fvecs and <prefix>_groundtruth.ivecs
// are read; the graph file is <prefix>.kgraph.graph, <prefix>.efanna.graph
// or <prefix>.nsg.graph (WEAVESS_GRAPH overrides).  K, L, ITER, S, R (and
// NSG's L_refine, R_refine, C) come from the per-dataset tables below (keyed
// by the prefix's file name) and can be overridden with WEAVESS_K, _L,
// _ITER, _S, _R, _L_REFINE, _R_REFINE, _C_REFINE; WEAVESS_SEED, WEAVESS_DTYPE (f32|bf16|fp8)
// and WEAVESS_EXACT (exact-kNN graph for "info") are read too.
// WEAVESS_ENTRY=centroid|subcentroid searches from the centroid entry point
// or the entry set eps_ instead of random entries.  WEAVESS_FORMAT=nsg|ssg
// reads and writes graph files headed by ep_ (NSG/Vamana layout) or eps_
// (SSG layout) instead of the plain layout; a build with WEAVESS_FORMAT=nsg
// finds the centroid entry point and writes it.  Without L_search, search
// tunes L to WEAVESS_ACC_SET (default 0.95) recall.  EFANNA builds start
// from InitRandom's graph (the reference's KD-tree initialisation is not
// built) and are written in the plain layout.  NSG builds run NN-Descent,
// then the NSG refine; Vamana builds run INIT_RAND and two Vamana passes
// (alpha 1, then 2).  Both are written in the NSG layout, and their search
// starts from the centroid entry point stored in the file.  DPG builds run
// NN-Descent, then the DPG refine, and search from random entries.  SSG
// builds run NN-Descent, then the SSG refine, and are written in the SSG
// layout with their roots eps_, which their search starts from (WEAVESS_A,
// WEAVESS_N_TRY override A = 60 and n_try = 10).
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>

#include "weavess/builder.h"
#include "weavess/component.h"

namespace {

struct KG {
    unsigned K, L, ITER, S, R;
};

// KGraph parameters per dataset (include/weavess/exp_data.h:57-98), plus
// the synthetic workloads of this project.
const std::map<std::string, KG> &table() {
    static const std::map<std::string, KG> t = {
        {"siftsmall", {25, 50, 6, 10, 100}}, {"sift1M", {90, 130, 12, 20, 50}},
        {"gist", {100, 120, 12, 25, 100}},   {"glove-100", {100, 150, 12, 35, 150}},
        {"c1", {50, 70, 10, 10, 100}},       {"c2", {64, 64, 12, 16, 64}},
    };
    return t;
}

// EFANNA's K, L, Iter, S, R per dataset (include/weavess/exp_data.h:315-340;
// its KD-tree nTrees / mLevel have no use here, the initial graph is
// InitRandom's).
const std::map<std::string, KG> &table_efanna() {
    static const std::map<std::string, KG> t = {
        {"siftsmall", {30, 50, 10, 15, 150}}, {"sift1M", {60, 70, 10, 15, 150}},
        {"gist", {100, 190, 7, 30, 50}},      {"glove-100", {100, 170, 7, 10, 100}},
        {"crawl", {100, 120, 8, 25, 100}},    {"msong", {50, 130, 7, 10, 150}},
        {"c1", {50, 70, 10, 10, 100}},        {"c2", {64, 64, 12, 16, 64}},
    };
    return t;
}

// NSG's K, L, Iter, S, R (its NN-Descent) and L_refine, R_refine, C per
// dataset (include/weavess/exp_data.h:104-140).
struct NS {
    KG kg;
    unsigned L_refine, R_refine, C;
};
const std::map<std::string, NS> &table_nsg() {
    static const std::map<std::string, NS> t = {
        {"siftsmall", {{25, 50, 6, 10, 100}, 100, 50, 500}}, {"sift1M", {{100, 120, 12, 25, 300}, 150, 30, 400}},
        {"gist", {{400, 430, 12, 10, 200}, 500, 20, 400}},   {"glove-100", {{400, 420, 12, 20, 300}, 150, 90, 600}},
        {"c1", {{50, 70, 10, 10, 100}, 40, 32, 200}},        {"c2", {{64, 64, 12, 16, 64}, 60, 32, 300}},
    };
    return t;
}

// Vamana's L_refine and R_refine per dataset; its INIT_RAND draws R
// neighbours a node ("L" = R) (include/weavess/exp_data.h:266-313).
const std::map<std::string, std::pair<unsigned, unsigned>> &table_vamana() {
    static const std::map<std::string, std::pair<unsigned, unsigned>> t = {
        {"siftsmall", {30, 20}}, {"sift1M", {70, 50}}, {"gist", {60, 50}}, {"glove-100", {120, 110}},
        {"c1", {40, 24}},        {"c2", {70, 50}},
    };
    return t;
}

// SSG's NN-Descent K, L, ITER, S, R and L_refine, R_refine per dataset
// (include/weavess/exp_data.h:159-214; A = 60, n_try = 10 for all).
const std::map<std::string, NS> &table_ssg() {
    static const std::map<std::string, NS> t = {
        {"siftsmall", {{25, 50, 6, 10, 100}, 100, 50, 0}},    {"sift1M", {{400, 420, 12, 20, 100}, 50, 20, 0}},
        {"gist", {{300, 330, 12, 20, 200}, 200, 40, 0}},      {"glove-100", {{300, 320, 12, 10, 200}, 150, 30, 0}},
        {"c1", {{50, 70, 10, 10, 100}, 50, 20, 0}},           {"c2", {{64, 64, 12, 16, 64}, 100, 32, 0}},
    };
    return t;
}

// DPG's NN-Descent K, L, ITER, S, R per dataset (include/weavess/exp_data.h:
// 215-264); the DPG refine keeps K / 2 a node.
const std::map<std::string, KG> &table_dpg() {
    static const std::map<std::string, KG> t = {
        {"siftsmall", {25, 50, 6, 10, 100}}, {"sift1M", {100, 100, 12, 25, 300}},
        {"gist", {100, 100, 12, 20, 100}},   {"glove-100", {100, 130, 12, 20, 100}},
        {"c1", {50, 70, 10, 10, 100}},       {"c2", {64, 64, 12, 16, 64}},
    };
    return t;
}

unsigned env_u(const char *name, unsigned dflt) {
    const char *v = std::getenv(name);
    return v && *v ? (unsigned)std::strtoul(v, nullptr, 10) : dflt;
}

}  // namespace

int main(int argc, char **argv) {
    const std::string alg = argc > 1 ? argv[1] : "";
    if (argc < 4 || argc > 5 ||
        (alg != "kgraph" && alg != "efanna" && alg != "nsg" && alg != "vamana" && alg != "dpg" && alg != "ssg")) {
        std::cout << "./weavess_main kgraph|efanna|nsg|vamana|dpg|ssg dataset_prefix build|search|info [L_search]"
                  << std::endl;
        return 1;
    }
    const bool efanna = alg == "efanna", nsg = alg == "nsg", vamana = alg == "vamana", dpg = alg == "dpg",
               ssg = alg == "ssg";
    const std::string prefix(argv[2]), exc(argv[3]);
    std::string name = prefix.substr(prefix.find_last_of('/') + 1);
    const auto &tab = efanna ? table_efanna() : table();
    KG kg = tab.count(name) ? tab.at(name) : tab.at("c1");
    NS ns = table_nsg().count(name) ? table_nsg().at(name) : table_nsg().at("c1");
    if (nsg) kg = ns.kg;
    if (dpg) kg = table_dpg().count(name) ? table_dpg().at(name) : table_dpg().at("c1");
    if (ssg) {
        ns = table_ssg().count(name) ? table_ssg().at(name) : table_ssg().at("c1");
        kg = ns.kg;
    }
    if (vamana) {
        const auto vm = table_vamana().count(name) ? table_vamana().at(name) : table_vamana().at("c1");
        ns.L_refine = vm.first;
        ns.R_refine = vm.second;
        kg.L = vm.second;  // INIT_RAND's neighbours a node
    }
    weavess::Parameters P;
    P.set<unsigned>("K", env_u("WEAVESS_K", kg.K));
    P.set<unsigned>("L", env_u("WEAVESS_L", kg.L));
    P.set<unsigned>("ITER", env_u("WEAVESS_ITER", kg.ITER));
    P.set<unsigned>("S", env_u("WEAVESS_S", kg.S));
    P.set<unsigned>("R", env_u("WEAVESS_R", kg.R));
    P.set<unsigned long long>("seed", env_u("WEAVESS_SEED", 0));
    if (const char *a = std::getenv("WEAVESS_ACC_SET")) P.set<std::string>("acc_set", a);
    else P.set<float>("acc_set", 0.95f);
    if (const char *e = std::getenv("WEAVESS_EXACT")) P.set<std::string>("exact_knng_path", e);
    P.set<unsigned>("L_refine", env_u("WEAVESS_L_REFINE", ns.L_refine));
    P.set<unsigned>("R_refine", env_u("WEAVESS_R_REFINE", ns.R_refine));
    P.set<unsigned>("C_refine", env_u("WEAVESS_C_REFINE", ns.C));
    P.set<float>("A", (float)env_u("WEAVESS_A", 60));  // include/weavess/exp_data.h:211-212
    P.set<unsigned>("n_try", env_u("WEAVESS_N_TRY", 10));
    const char *gf = std::getenv("WEAVESS_GRAPH");
    const std::string graph = gf ? gf
                                 : prefix + (efanna   ? ".efanna.graph"
                                             : nsg    ? ".nsg.graph"
                                             : vamana ? ".vamana.graph"
                                             : dpg    ? ".dpg.graph"
                                             : ssg    ? ".ssg.graph"
                                                      : ".kgraph.graph");
    const std::string base = prefix + "_base.fvecs", query = prefix + "_query.fvecs",
                      ground = prefix + "_groundtruth.ivecs";
    const std::string dtype = std::getenv("WEAVESS_DTYPE") ? std::getenv("WEAVESS_DTYPE") : "f32";
    const weavess::TYPE refine =
        efanna ? (dtype == "bf16"  ? weavess::REFINE_EFANNA_HIP_BF16
                  : dtype == "fp8" ? weavess::REFINE_EFANNA_HIP_FP8
                                   : weavess::REFINE_EFANNA)
               : (dtype == "bf16"  ? weavess::REFINE_NN_DESCENT_HIP_BF16
                  : dtype == "fp8" ? weavess::REFINE_NN_DESCENT_HIP_FP8
                                   : weavess::REFINE_NN_DESCENT_HIP);
    const std::string ent =
        std::getenv("WEAVESS_ENTRY") ? std::getenv("WEAVESS_ENTRY")
        : (nsg || vamana)            ? "centroid"
        : ssg                        ? "subcentroid"
                                     : "rand";
    const weavess::TYPE entry = ent == "centroid"      ? weavess::SEARCH_ENTRY_CENTROID
                                : ent == "subcentroid" ? weavess::SEARCH_ENTRY_SUB_CENTROID
                                                       : weavess::SEARCH_ENTRY_RAND;
    const std::string fmt =
        std::getenv("WEAVESS_FORMAT") ? std::getenv("WEAVESS_FORMAT")
        : (nsg || vamana)             ? "nsg"
        : ssg                         ? "ssg"
                                      : "kgraph";
    const weavess::TYPE gtype = fmt == "nsg" ? weavess::INDEX_NSG : fmt == "ssg" ? weavess::INDEX_SSG
                                                                                 : weavess::INDEX_KGRAPH;
    try {
        weavess::IndexBuilder builder(8);
        if (exc == "build") {
            if (gtype == weavess::INDEX_SSG && !ssg) throw std::invalid_argument("SSG files are written by ssg builds");
            builder.load(base.c_str(), query.c_str(), ground.c_str(), P)
                ->init(vamana ? weavess::INIT_RAND : weavess::INIT_RANDOM);
            std::cout << "Init cost: " << builder.GetBuildTime().count() << std::endl;
            if (vamana)  // test/main.cpp:168-174
                builder.refine(weavess::REFINE_VAMANA)->refine(weavess::REFINE_VAMANA);
            else  // WEAVESS_DEBUG=1: the per-round graph-recall printout of the refine's debug path
                builder.refine(refine, std::getenv("WEAVESS_DEBUG") && std::string(std::getenv("WEAVESS_DEBUG")) == "1");
            if (nsg || dpg || ssg) {  // test/main.cpp:76-81, :106-112, :137-142
                std::cout << "Init cost: " << builder.GetBuildTime().count() << std::endl;
                builder.refine(nsg ? weavess::REFINE_NSG : dpg ? weavess::REFINE_DPG : weavess::REFINE_SSG);
            } else if (gtype == weavess::INDEX_NSG) {
                weavess::ComponentRefineEntryCentroidHIP(builder.index()).EntryInner();
            }
            builder.save_graph(gtype, graph.c_str());
            std::cout << "Build cost: " << builder.GetBuildTime().count() << std::endl;
        } else if (exc == "search") {
            builder.load(base.c_str(), query.c_str(), ground.c_str(), P)->load_graph(gtype, graph.c_str());
            if (argc == 5) {
                builder.index()->getParam().set<unsigned>("L_search", (unsigned)std::atoi(argv[4]));
                builder.search(entry, weavess::ROUTER_GREEDY_HIP, weavess::L_SEARCH_ASSIGN);
            } else {
                builder.search(entry, weavess::ROUTER_GREEDY_HIP, weavess::L_SEARCH_SET_RECALL);
            }
            std::cout << "QPS: " << builder.last_qps() << " recall: " << builder.last_recall()
                      << " L: " << builder.last_L() << std::endl;
            builder.peak_memory_footprint();
        } else if (exc == "info") {
            builder.load(base.c_str(), query.c_str(), ground.c_str(), P)
                ->load_graph(gtype, graph.c_str())
                ->print_index_info(gtype);
        } else {
            std::cout << "exc_type input error!" << std::endl;
            return 1;
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
