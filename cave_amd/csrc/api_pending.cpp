// api_pending.cpp -- entry points whose engines are not built yet.
#include "common.h"

using namespace cave;

extern "C" {

int cave_search_create(const CaveSearchSpec *, CaveSearch **out)
{
    if (out) *out = nullptr;
    return fail(CAVE_E_UNSUPPORTED, "search engine not built");
}
int cave_search_step(CaveSearch *, uint64_t) { return fail(CAVE_E_UNSUPPORTED, "search engine not built"); }
int cave_search_is_running(const CaveSearch *) { return 0; }
int cave_search_progress(const CaveSearch *, CaveSearchProgress *) { return fail(CAVE_E_UNSUPPORTED, "not built"); }
int cave_search_drain(CaveSearch *, CaveSearchResult *, size_t, size_t *) { return fail(CAVE_E_UNSUPPORTED, "not built"); }
int cave_search_write_summary(CaveSearch *, char *, size_t) { return fail(CAVE_E_UNSUPPORTED, "not built"); }
int cave_search_destroy(CaveSearch *) { return CAVE_OK; }

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
