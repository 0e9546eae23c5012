// refcount.hpp -- shared ownership of a context by the engines made on it.
//
// The reference holds its engines through Reference::To (Fold.h:243) and
// gives every engine the stream of its thread; an engine may outlive the
// code that made the stream.  Here the context is counted: the caller's
// handle holds one reference and every engine made on the context holds
// another, so destroying the context before its engines only drops the
// caller's reference, and the context is freed when the last engine is
// destroyed.  Header-only and free of HIP so that the rule can be checked
// on the CPU under AddressSanitizer.
#pragma once
#include <atomic>

namespace dspx {

struct RefCount {
  std::atomic<int> n{1};  // the creator's reference
};

// One reference on an object T that embeds a RefCount `refs`; FREE(T*) runs
// when the last reference is dropped.
template <class T, void (*FREE)(T*)> class Ref {
 public:
  Ref() = default;
  Ref(const Ref&) = delete;
  Ref& operator=(const Ref&) = delete;
  ~Ref() { reset(nullptr); }
  void reset(T* x)
  {
    if (x) x->refs.n.fetch_add(1, std::memory_order_relaxed);
    T* old = p_;
    p_ = x;
    if (old) drop(old);
  }
  T* get() const { return p_; }
  // Drops one reference of x (the creator's, for the owner's destroy call).
  static void drop(T* x)
  {
    if (x->refs.n.fetch_sub(1, std::memory_order_acq_rel) == 1) FREE(x);
  }

 private:
  T* p_ = nullptr;
};

}  // namespace dspx
