// api_pending.cpp -- entry points whose engines are not built yet.
#include "common.h"

using namespace cave;

extern "C" {

int cave_serve_create(const CaveServeOptions *, CaveServer **out)
{
    if (out) *out = nullptr;
    return fail(CAVE_E_UNSUPPORTED, "server not built");
}
int cave_serve_submit(CaveServer *, const CaveSimSpec *, uint64_t) { return fail(CAVE_E_UNSUPPORTED, "not built"); }
int cave_serve_poll(CaveServer *, CaveServeCompletion *, size_t, size_t *) { return fail(CAVE_E_UNSUPPORTED, "not built"); }
int cave_serve_drain(CaveServer *) { return fail(CAVE_E_UNSUPPORTED, "not built"); }
int cave_serve_stats(const CaveServer *, uint64_t *, uint64_t *, uint64_t *, uint64_t *) { return fail(CAVE_E_UNSUPPORTED, "not built"); }
int cave_serve_destroy(CaveServer *) { return CAVE_OK; }

}
