// Test infrastructure only: a thin C-ABI driver around the *reference* nmslib
// HNSW sources, compiled from /root/reference by oracle/Makefile into
// oracle/_ref/.  It is the checker for golden generation and the optional
// "reference" CPU baseline; nothing in the product links it.
//
// It replays what python_bindings/nmslib.cc does for one dense float index:
//   createIndex      nmslib.cc:89-96       (Hnsw::CreateIndex, hnsw.cc:183)
//   saveIndex        nmslib.cc:115-125
//   loadIndex        nmslib.cc:98-113      (LoadIndex + ResetQueryTimeParams)
//   setQueryTimeParams nmslib.cc:599-607
//   knnQueryBatch    nmslib.cc:142-184     (ParallelFor over Index::Search,
//                                            heap popped into ascending arrays)
//   seq_search       method/seqsearch.cc:143-177 (brute force, single thread)
#include <cstdint>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "init.h"
#include "knnquery.h"
#include "knnqueue.h"
#include "method/hnsw.h"
#include "method/seqsearch.h"
#include "object.h"
#include "params.h"
#include "space/space_lp.h"
#include "space/space_scalar.h"
#include "space/space_vector.h"
#include "thread_pool.h"

using namespace similarity;

namespace {

thread_local std::string g_err;

struct RefIndex {
  std::unique_ptr<Space<float>> space;
  ObjectVector data;
  std::unique_ptr<Index<float>> index;
  int dim = 0;
  ~RefIndex() {
    index.reset();
    for (auto* o : data) delete o;
  }
};

Space<float>* make_space(const char* name) {
  std::string s(name);
  if (s == "l2") return new SpaceLp<float>(2);
  if (s == "l1") return new SpaceLp<float>(1);
  if (s == "linf") return new SpaceLp<float>(-1);
  if (s == "cosinesimil") return new SpaceCosineSimilarity<float>();
  if (s == "negdotprod") return new SpaceNegativeScalarProduct<float>();
  throw std::runtime_error("ref driver: unsupported space " + s);
}

std::vector<std::string> split_params(const char* p) {
  // comma separated "k=v" list
  std::vector<std::string> out;
  if (!p || !*p) return out;
  std::string s(p), cur;
  for (char c : s) {
    if (c == ',') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

bool g_inited = false;
void ensure_init() {
  if (!g_inited) {
    initLibrary(0, LIB_LOGNONE, nullptr);
    g_inited = true;
  }
}

}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

void* ref_create(const char* space, const float* data, int64_t n, int dim) {
  try {
    ensure_init();
    auto* r = new RefIndex();
    r->space.reset(make_space(space));
    r->dim = dim;
    auto* vs = dynamic_cast<VectorSpace<float>*>(r->space.get());
    std::vector<float> v(dim);
    for (int64_t i = 0; i < n; ++i) {
      std::memcpy(v.data(), data + i * dim, sizeof(float) * dim);
      r->data.push_back(vs->CreateObjFromVect((IdType)i, -1, v));
    }
    return r;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

int ref_create_index(void* h, const char* params) {
  try {
    auto* r = static_cast<RefIndex*>(h);
    r->index.reset(new Hnsw<float>(false, *r->space, r->data));
    r->index->CreateIndex(AnyParams(split_params(params)));
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

int ref_save_index(void* h, const char* path) {
  try {
    static_cast<RefIndex*>(h)->index->SaveIndex(path);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void* ref_load(const char* space, const char* path, int dim) {
  try {
    ensure_init();
    auto* r = new RefIndex();
    r->space.reset(make_space(space));
    r->dim = dim;
    r->index.reset(new Hnsw<float>(false, *r->space, r->data));
    r->index->LoadIndex(path);
    r->index->ResetQueryTimeParams();
    return r;
  } catch (const std::exception& e) {
    g_err = e.what();
    return nullptr;
  }
}

int ref_set_query_params(void* h, const char* params) {
  try {
    static_cast<RefIndex*>(h)->index->SetQueryTimeParams(AnyParams(split_params(params)));
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Same flow as IndexWrapper::knnQueryBatch; results are written row-major
// [nq][k] ascending, counts[q] = number of results found.
int ref_knn_query_batch(void* h, const float* q, int64_t nq, int dim, int k,
                        int num_threads, int32_t* ids, float* dists, int32_t* counts) {
  try {
    auto* r = static_cast<RefIndex*>(h);
    auto* vs = dynamic_cast<VectorSpace<float>*>(r->space.get());
    ObjectVector queries;
    std::vector<float> v(dim);
    for (int64_t i = 0; i < nq; ++i) {
      std::memcpy(v.data(), q + i * dim, sizeof(float) * dim);
      queries.push_back(vs->CreateObjFromVect(0, -1, v));
    }
    std::vector<std::unique_ptr<KNNQueue<float>>> results(nq);
    ParallelFor(0, nq, num_threads, [&](size_t qi, size_t) {
      KNNQuery<float> knn(*r->space, queries[qi], k);
      r->index->Search(&knn, -1);
      results[qi].reset(knn.Result()->Clone());
    });
    for (auto* o : queries) delete o;
    for (int64_t i = 0; i < nq; ++i) {
      KNNQueue<float>* res = results[i].get();
      size_t size = res->Size();
      counts[i] = (int32_t)size;
      for (int j = (int)size; j < k; ++j) {
        ids[i * k + j] = -1;
        dists[i * k + j] = 0.f;
      }
      while (!res->Empty() && size > 0) {
        size -= 1;
        ids[i * k + size] = res->TopObject()->id();
        dists[i * k + size] = res->TopDistance();
        res->Pop();
      }
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

// Brute force through the reference's seq_search method over the objects
// held by a ref_create() handle (distances as the space defines them: l2 is
// the square root of the squared sum, spaces/space_lp.h:57-58).
int ref_seq_search(void* h, const float* q, int64_t nq, int dim, int k,
                   int32_t* ids, float* dists, int32_t* counts) {
  try {
    auto* r = static_cast<RefIndex*>(h);
    auto* vs = dynamic_cast<VectorSpace<float>*>(r->space.get());
    SeqSearch<float> seq(*r->space, r->data);
    seq.CreateIndex(AnyParams());
    std::vector<float> v(dim);
    for (int64_t i = 0; i < nq; ++i) {
      std::memcpy(v.data(), q + i * dim, sizeof(float) * dim);
      std::unique_ptr<Object> qo(vs->CreateObjFromVect(0, -1, v));
      KNNQuery<float> knn(*r->space, qo.get(), k);
      seq.Search(&knn, -1);
      std::unique_ptr<KNNQueue<float>> res(knn.Result()->Clone());
      size_t size = res->Size();
      counts[i] = (int32_t)size;
      while (!res->Empty() && size > 0) {
        size -= 1;
        ids[i * k + size] = res->TopObject()->id();
        dists[i * k + size] = res->TopDistance();
        res->Pop();
      }
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  }
}

void ref_free(void* h) { delete static_cast<RefIndex*>(h); }

}  // extern "C"
