// Per-pointer caches of derived device structures are keyed on the caller's device pointer;
// fe_gpu_runtime_free calls forget_pointer so a freed (and possibly reused) address never
// matches a stale entry.
#include <mutex>
#include <unordered_map>
#include <vector>

namespace fehip {

namespace {
std::mutex g_reg_mutex;
std::unordered_map<const void*, std::vector<void (*)(const void*)>>& hooks() {
    static std::unordered_map<const void*, std::vector<void (*)(const void*)>> h;
    return h;
}
}  // namespace

void register_pointer_hook(const void* ptr, void (*fn)(const void*)) {
    std::lock_guard<std::mutex> lock(g_reg_mutex);
    hooks()[ptr].push_back(fn);
}

void forget_pointer(const void* ptr) {
    std::vector<void (*)(const void*)> fns;
    {
        std::lock_guard<std::mutex> lock(g_reg_mutex);
        auto it = hooks().find(ptr);
        if (it == hooks().end()) return;
        fns.swap(it->second);
        hooks().erase(it);
    }
    for (auto fn : fns) fn(ptr);
}

}  // namespace fehip
