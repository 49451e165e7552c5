#include "GpuContext.h"

#include "../../include/sfp.h"

#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>

namespace Pivot::detail {

namespace {
using Key = std::tuple<int, int, int, double, double, double, double>;
struct Registry {
    std::mutex lock;
    std::map<Key, sfp_context> contexts;
    ~Registry() {
        for (auto &kv : contexts) sfp_destroy(kv.second);
    }
};
Registry &registry() {
    static Registry r;
    return r;
}
}  // namespace

[[noreturn]] void Fail(sfp_context_s *ctx, int status, char const *what) {
    std::fprintf(stderr, "%s failed: %s (%s)\n", what, sfp_status_string((sfp_status)status),
                 sfp_last_error(ctx));
    std::exit(EXIT_FAILURE);
}

sfp_context_s *SharedContext(Vector3i const &n, double spacing, Vector3d const &o) {
    Registry &r = registry();
    std::lock_guard<std::mutex> g(r.lock);
    Key const key{n[0], n[1], n[2], spacing, o[0], o[1], o[2]};
    auto it = r.contexts.find(key);
    if (it != r.contexts.end()) return it->second;
    double const origin[3] = {o[0], o[1], o[2]};
    sfp_context ctx = nullptr;
    sfp_status const st = sfp_create(&ctx, n[0], n[1], n[2], spacing, origin, 0);
    if (st != SFP_OK) Fail(nullptr, st, "sfp_create");
    r.contexts.emplace(key, ctx);
    return ctx;
}

}  // namespace Pivot::detail
