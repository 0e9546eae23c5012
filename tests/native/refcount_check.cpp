// The context's reference count (csrc/refcount.hpp) with the owner's handle
// dropped before its engines: built under AddressSanitizer by
// tests/test_refcount.py, which fails if any access below touches freed
// memory or the object is freed a number of times other than once.
#include <cstdio>

#include "refcount.hpp"

struct Fake {
  dspx::RefCount refs;
  int payload = 42;
};

static int freed = 0;
static void free_fake(Fake* f)
{
  freed++;
  delete f;
}
using FakeRef = dspx::Ref<Fake, free_fake>;

struct Engine {
  FakeRef hold;  // first member: dropped after the engine's own state
  int* state = new int(7);
  ~Engine() { delete state; }
};

int main()
{
  Fake* ctx = new Fake;  // the creator's reference
  Engine* a = new Engine;
  Engine* b = new Engine;
  a->hold.reset(ctx);
  b->hold.reset(ctx);
  FakeRef::drop(ctx);  // dspx_ctx_destroy before the engines
  if (freed != 0) return 1;
  int sum = a->hold.get()->payload + b->hold.get()->payload;  // still alive
  delete a;
  if (freed != 0) return 2;
  sum += b->hold.get()->payload;
  delete b;  // the last reference
  if (freed != 1) return 3;
  // an engine moved to another context drops the first one
  Fake* c1 = new Fake;
  Fake* c2 = new Fake;
  Engine* e = new Engine;
  e->hold.reset(c1);
  FakeRef::drop(c1);
  e->hold.reset(c2);  // c1 freed here
  if (freed != 2) return 4;
  FakeRef::drop(c2);
  delete e;
  if (freed != 3) return 5;
  std::printf("refcount ok %d\n", sum);
  return 0;
}
