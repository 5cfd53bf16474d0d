// dk_prof.hip -- per-kernel timing with HIP events on the context's stream.
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "dk_internal.hpp"

struct Prof {
  bool on = false;
  struct Pending {std::string label; hipEvent_t a, b;};
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  std::vector<std::string> order;
  std::map<std::string, std::pair<double, long long>> acc;

  hipEvent_t Get() {
    if (!pool.empty()) {hipEvent_t e = pool.back(); pool.pop_back(); return e;}
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
  }
  void Collect(hipStream_t s) {
    if (pending.empty()) return;
    (void)hipStreamSynchronize(s);
    for (auto &p : pending) {
      float ms = 0.0f;
      if (p.a && p.b && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        auto it = acc.find(p.label);
        if (it == acc.end()) {order.push_back(p.label); it = acc.emplace(p.label,
                                                                 std::make_pair(0.0, 0LL)).first;}
        it->second.first += ms;
        it->second.second += 1;
      }
      if (p.a) pool.push_back(p.a);
      if (p.b) pool.push_back(p.b);
    }
    pending.clear();
  }
};

namespace dk {

bool ProfOn(const dk_ctx *c) { return c->prof != nullptr && c->prof->on; }

int ProfBegin(dk_ctx *c, const char *label) {
  Prof *p = c->prof;
  if (p == nullptr || !p->on) return -1;
  if (p->pending.size() >= 65536) p->Collect(c->stream);
  Prof::Pending e{label, p->Get(), p->Get()};
  if (e.a) (void)hipEventRecord(e.a, c->stream);
  p->pending.push_back(e);
  return static_cast<int>(p->pending.size()) - 1;
}

void ProfEnd(dk_ctx *c, int token) {
  Prof *p = c->prof;
  if (p == nullptr || !p->on || token < 0 || token >= static_cast<int>(p->pending.size()))
    return;
  if (p->pending[token].b) (void)hipEventRecord(p->pending[token].b, c->stream);
}

void ProfDestroy(dk_ctx *c) {
  Prof *p = c->prof;
  if (p == nullptr) return;
  p->Collect(c->stream);
  for (hipEvent_t e : p->pool) (void)hipEventDestroy(e);
  delete p;
  c->prof = nullptr;
}

}  // namespace dk

extern "C" {

int dk_prof_enable(dk_ctx *c, int32_t on) {
  if (c == nullptr) return DK_EINVAL;
  if (c->prof == nullptr) c->prof = new Prof();
  c->prof->Collect(c->stream);
  c->prof->on = (on != 0);
  return DK_OK;
}

int dk_prof_reset(dk_ctx *c) {
  if (c == nullptr) return DK_EINVAL;
  if (c->prof == nullptr) return DK_OK;
  c->prof->Collect(c->stream);
  c->prof->acc.clear();
  c->prof->order.clear();
  return DK_OK;
}

int dk_prof_read(dk_ctx *c, int32_t idx, char *label, int32_t label_len, double *ms,
                 int64_t *count) {
  if (c == nullptr || c->prof == nullptr) return DK_EINVAL;
  c->prof->Collect(c->stream);
  if (idx < 0 || idx >= static_cast<int32_t>(c->prof->order.size())) return DK_EINVAL;
  const std::string &l = c->prof->order[idx];
  const auto &v = c->prof->acc[l];
  if (label && label_len > 0) {
    std::strncpy(label, l.c_str(), label_len - 1);
    label[label_len - 1] = '\0';
  }
  if (ms) *ms = v.first;
  if (count) *count = v.second;
  return DK_OK;
}

}  // extern "C"
