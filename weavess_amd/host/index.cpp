// index.cpp -- Index: data ownership and the device handles of the HIP stages.
#include "weavess/index.h"

#include <algorithm>

namespace weavess {

Index::Index(int device) : device_(device) {}

Index::~Index() {
    if (dix_) wv_index_destroy(dix_);
    if (ctx_) wv_ctx_destroy(ctx_);
}

void Index::setBase(std::vector<float> v, unsigned n, unsigned d) {
    base_ = std::move(v);
    base_len_ = n;
    base_dim_ = d;
    if (dix_) {
        wv_index_destroy(dix_);
        dix_ = nullptr;
    }
    load_graph_dirty_ = true;
}
void Index::setQuery(std::vector<float> v, unsigned n, unsigned d) {
    query_ = std::move(v);
    query_len_ = n;
    query_dim_ = d;
}
void Index::setGround(std::vector<unsigned> v, unsigned n, unsigned d) {
    ground_ = std::move(v);
    ground_len_ = n;
    ground_dim_ = d;
}

void Index::check(int rc) const {
    if (rc != WV_OK) throw HipError(rc, std::string("weavess_hip: ") + wv_last_error(ctx_));
}

wv_ctx *Index::ctx() {
    if (!ctx_) {
        int rc = wv_ctx_create(device_, &ctx_);
        if (rc != WV_OK) {
            ctx_ = nullptr;
            throw HipError(rc, "wv_ctx_create failed (no usable GPU?)");
        }
    }
    return ctx_;
}

wv_index *Index::deviceIndex() {
    if (!dix_) {
        if (base_len_ == 0) throw std::invalid_argument("no base data loaded");
        check(wv_index_create(ctx(), base_.data(), base_len_, base_dim_, &dix_));
        load_graph_dirty_ = true;
    }
    return dix_;
}

void Index::uploadLoadGraph() {
    wv_index *ix = deviceIndex();
    if (!load_graph_dirty_) return;
    if (load_graph_.size() != base_len_)
        throw std::invalid_argument("graph has " + std::to_string(load_graph_.size()) + " nodes, base has " +
                                    std::to_string(base_len_));
    unsigned dmax = 1;
    for (const auto &r : load_graph_) dmax = std::max<unsigned>(dmax, (unsigned)r.size());
    std::vector<uint32_t> ids((size_t)base_len_ * dmax, 0xFFFFFFFFu), deg(base_len_);
    for (unsigned i = 0; i < base_len_; ++i) {
        deg[i] = (uint32_t)load_graph_[i].size();
        std::copy(load_graph_[i].begin(), load_graph_[i].end(), ids.begin() + (size_t)i * dmax);
    }
    check(wv_index_set_graph(ix, ids.data(), deg.data(), dmax));
    load_graph_dirty_ = false;
}

}  // namespace weavess
