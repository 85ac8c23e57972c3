// Stub: one-rank MPI context, shared_object and the MPI_Bcast the HF hook calls.
#pragma once
#include <utility>

typedef int MPI_Comm;
typedef int MPI_Datatype;
#define MPI_DOUBLE 1
inline int MPI_Bcast(void*, int, MPI_Datatype, int, MPI_Comm) { return 0; }

namespace green::utils {
  struct mpi_context {
    int      global_rank = 0, node_rank = 0, internode_size = 1;
    MPI_Comm global = 0;
  };
  inline const mpi_context& context() {
    static mpi_context c;
    return c;
  }
  template <typename T>
  class shared_object {
  public:
    explicit shared_object(T obj) : obj_(std::move(obj)) {}
    void fence() {}
    T&   object() { return obj_; }

  private:
    T obj_;
  };
}  // namespace green::utils
