// lpa/graph.h -- host graph containers with the reference's constructors and
// members (include/graph.h:17-115).  The CGR offsets are 64-bit.
#pragma once
#include <cstdint>
#include <iostream>
#include <memory>
#include <vector>

namespace lpa_api {

template <typename T, typename... Args> std::unique_ptr<T> make_unique_ptr(Args &&...args) {
    return std::unique_ptr<T>(new T(std::forward<Args>(args)...));
}

template <typename V, typename E> class Graph {
  public:
    Graph() = default;
    Graph(V _n, E _m) : n(_n), m(_m) {}
    virtual ~Graph() = default;
    V n;   // vertices
    E m;   // arcs
};

template <typename V, typename E> class CSRGraph : public Graph<V, E> {
  public:
    CSRGraph() = default;
    CSRGraph(std::vector<V> _neighbors, std::vector<E> _offsets)
        : Graph<V, E>((V)(_offsets.size() - 1), (E)_neighbors.size()), neighbors(std::move(_neighbors)),
          offsets(std::move(_offsets)) {}
    std::vector<V> neighbors;
    std::vector<E> offsets;
};

// CGR part (words, 64-bit start offsets) plus the CSR of the small vertices
template <typename V, typename E> class CGRGraph : public Graph<V, E> {
  public:
    CGRGraph() = default;
    CGRGraph(std::vector<V> &_csr_neighbors, std::vector<E> &_csr_offsets, std::vector<V> &_cgr_neighbors,
             std::vector<uint64_t> &_cgr_offsets, std::vector<V> &_csr_small_neighbors,
             std::vector<E> &_csr_small_offsets, std::vector<V> &_vertices, int sns)
        : Graph<V, E>((V)(_csr_offsets.size() - 1), (E)_csr_neighbors.size()), csr_neighbors(_csr_neighbors),
          csr_offsets(_csr_offsets), cgr_neighbors(_cgr_neighbors), cgr_offsets(_cgr_offsets),
          csr_small_neighbors(_csr_small_neighbors), csr_small_offsets(_csr_small_offsets),
          vertices(_vertices), small_node_thresh(sns) {
        n_small = _csr_small_offsets.empty() ? 0 : (int)(_csr_small_offsets.size() - 1);
        m_small = (int)_csr_small_neighbors.size();
        std::cout << "CGR mixed graph: n " << this->n << " m " << this->m << " compressed "
                  << (_cgr_offsets.empty() ? 0 : _cgr_offsets.size() - 1) << " small " << n_small << "\n";
    }
    std::vector<V> csr_neighbors;
    std::vector<E> csr_offsets;
    std::vector<V> cgr_neighbors;
    std::vector<uint64_t> cgr_offsets;
    std::vector<V> csr_small_neighbors;
    std::vector<E> csr_small_offsets;
    std::vector<V> vertices;
    int small_node_thresh = 0;
    int n_small = 0;
    int m_small = 0;
};

// every vertex compressed (the CGLP method's input)
template <typename V, typename E> class CGRGraphPure : public Graph<V, E> {
  public:
    CGRGraphPure() = default;
    CGRGraphPure(std::vector<V> &_csr_neighbors, std::vector<E> &_csr_offsets, std::vector<V> &_cgr_neighbors,
                 std::vector<uint64_t> &_cgr_offsets, std::vector<V> &_vertices)
        : Graph<V, E>((V)(_csr_offsets.size() - 1), (E)_csr_neighbors.size()), csr_neighbors(_csr_neighbors),
          csr_offsets(_csr_offsets), cgr_neighbors(_cgr_neighbors), cgr_offsets(_cgr_offsets),
          vertices(_vertices) {
        std::cout << "CGR pure graph: n " << this->n << " m " << this->m << "\n";
    }
    std::vector<V> csr_neighbors;
    std::vector<E> csr_offsets;
    std::vector<V> cgr_neighbors;
    std::vector<uint64_t> cgr_offsets;
    std::vector<V> vertices;
};

}  // namespace lpa_api
