// fit_host.cpp - trajectory recording (ghip_record*) and the field replay
// with its adjoint (ghip_fit_*), over the kernels of adjoint.hip.
//
// Forward: the fields of every `K`-th step are kept (checkpoints). Backward:
// the segments are taken last first; each is recomputed from its checkpoint
// keeping every step's starting fields, then each step, last first, is
// recomputed keeping its substep states and run backwards: the transposed
// exchange, then the transposed substeps. Memory: checkpoints (steps / K),
// K step starts and substeps + 1 substep states, each n_species * n_sites
// f32, plus the one-byte owner maps of the trajectory.
//
// On a 2D lattice whose substep count fits (fit_fused_fits) a step is one
// launch each way, the substep states kept in LDS; elsewhere, or with
// GLAZIER_FIT_PER_SUBSTEP=1 in the environment at creation, one launch per
// substep and exchange. Both give the same fields and adjoint fields bit for
// bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/glazier_hip.h"
#include "fit.h"
#include "sim.h"

using namespace ghip;

struct ghip_fit {
    ghip_model m{};
    FitGrid g{};
    uint32_t S = 0;              // species
    int max_sub = 0;
    uint64_t steps = 0;
    uint32_t K = 20;
    uint8_t *types = nullptr;    // steps * n
    float *ckpt = nullptr;       // ceil(steps / K) * S * n
    float *seg = nullptr;        // K * S * n
    float *sub = nullptr;        // max_sub * S * n (substep states 1..max_sub)
    float *work[2] = {nullptr, nullptr};
    float *lam[2] = {nullptr, nullptr};
    double *part = nullptr;      // kPartWords per (species, workgroup)
    double *grad = nullptr;      // kPartWords per species
    hipStream_t own = nullptr;
    FitCoef coef{};
    bool have_forward = false;
    bool fused = false;          // one launch per step (2D)

    size_t field_words() const { return (size_t)S * g.n; }
    uint32_t part_blocks() const { return fused ? fit_tiles(g) : fit_blocks(g); }

    ~ghip_fit() {
        if (own) (void)hipStreamSynchronize(own);
        void *bufs[] = {types, ckpt, seg, sub, work[0], work[1], lam[0], lam[1], part, grad};
        for (void *p : bufs)
            if (p) (void)hipFree(p);
        if (own) (void)hipStreamDestroy(own);
    }
};

namespace {

FitCoef coef_of(const ghip_fit *f, const ghip_field_params &p) {
    FitCoef c{};
    for (uint32_t k = 0; k < f->S; ++k) {
        const int sub = f->m.species[k].substeps;
        c.sub[k] = sub;
        // as the engine converts its model (engine.cpp, ghip_create)
        c.d[k] = (float)(p.diffusion[k] / (double)sub);
        c.kd[k] = (float)(p.decay[k] / (double)sub);
        for (uint32_t t = 0; t < f->m.n_types; ++t) {
            c.sec[t][k] = (float)p.secretion[t][k];
            c.up[t][k] = (float)p.uptake[t][k];
        }
    }
    return c;
}

// One whole step m from `src` into `dst` (neither a work buffer), one
// launch per substep and exchange.
ghip_status forward_step(ghip_fit *f, const float *src, float *dst, uint64_t m, hipStream_t st) {
    const float *a = src;
    int w = 0;
    for (int s = 1; s <= f->max_sub; ++s) {
        HIP_TRY(launch_fit_substep(f->g, f->coef, f->S, a, f->work[w], s, st));
        a = f->work[w];
        w ^= 1;
    }
    HIP_TRY(launch_fit_exchange(f->g, f->coef, f->S, f->types + m * f->g.n, a, dst, st));
    return GHIP_OK;
}

// Steps m0 .. m0 + count - 1 from `src`: the state after step m0 + q goes
// to out_base + q * fw, the last one to `last`. No destination is `src` or a
// work buffer.
ghip_status forward_steps(ghip_fit *f, const float *src, uint64_t m0, uint64_t count, float *out_base,
                          float *last, hipStream_t st) {
    const size_t fw = f->field_words();
    auto dst_of = [&](uint64_t q) { return q + 1 == count ? last : out_base + q * fw; };
    if (!f->fused) {
        const float *cur = src;
        for (uint64_t q = 0; q < count; ++q) {
            ghip_status s = forward_step(f, cur, dst_of(q), m0 + q, st);
            if (s != GHIP_OK) return s;
            cur = dst_of(q);
        }
        return GHIP_OK;
    }
    const uint64_t per = (uint64_t)fit_steps_per_launch();
    const float *cur = src;
    for (uint64_t q = 0; q < count; q += per) {
        const uint64_t n = std::min(per, count - q);
        HIP_TRY(launch_fit_steps_fused(f->g, f->coef, f->S, f->max_sub, f->types + (m0 + q) * f->g.n, (int)n,
                                       cur, out_base + q * fw, fw, dst_of(q + n - 1), st));
        cur = dst_of(q + n - 1);
    }
    return GHIP_OK;
}

}  // namespace

extern "C" {

ghip_status ghip_record(ghip_sim *s, uint64_t max_steps) {
    if (!s) return fail(GHIP_BAD_ARG, "null handle");
    HIP_TRY(hipStreamSynchronize(s->stream));
    if (s->traj) HIP_TRY(hipFree(s->traj));
    s->traj = nullptr;
    s->traj_cap = s->traj_len = 0;
    if (max_steps == 0) return GHIP_OK;
    HIP_TRY(hipMalloc(&s->traj, max_steps * s->N));
    s->traj_cap = max_steps;
    return GHIP_OK;
}

ghip_status ghip_recorded_steps(const ghip_sim *s, uint64_t *out) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    *out = s->traj_len;
    return GHIP_OK;
}

ghip_status ghip_recorded_types(const ghip_sim *s, uint64_t step, uint8_t *out, size_t n) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    if (step >= s->traj_len) return fail(GHIP_BAD_ARG, "no such recorded step");
    if (n != s->N) return fail(GHIP_BAD_ARG, "the buffer does not match the lattice");
    HIP_TRY(hipStreamSynchronize(s->stream));
    HIP_TRY(hipMemcpy(out, s->traj + step * s->N, n, hipMemcpyDeviceToHost));
    return GHIP_OK;
}

ghip_status ghip_fit_create(const ghip_sim *s, uint32_t checkpoint_every, ghip_fit **out) {
    if (!s || !out) return fail(GHIP_BAD_ARG, "null argument");
    *out = nullptr;
    if (!s->traj || s->traj_len == 0) return fail(GHIP_BAD_ARG, "the lattice has recorded no steps");
    if (s->m.n_species == 0) return fail(GHIP_BAD_ARG, "the model has no fields to replay");
    if (checkpoint_every == 0) return fail(GHIP_BAD_ARG, "the checkpoint interval must be positive");
    if (s->N >= 0xFFFFFFFFull) return fail(GHIP_UNSUPPORTED, "the lattice exceeds 2^32 sites");
    ghip_fit *f = new ghip_fit();
    auto cleanup = [&](ghip_status st) {
        delete f;
        return st;
    };
#define FIT_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return cleanup(fail(GHIP_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e))); \
    } while (0)
    f->m = s->m;
    for (uint32_t k = 0; k < s->m.n_species; ++k) {
        f->m.species[k].substeps = s->substeps[k];  // the counts the engine settled on
        f->max_sub = std::max(f->max_sub, s->substeps[k]);
    }
    f->g.w = s->W;
    f->g.h = s->H;
    f->g.d = s->D;
    f->g.n = (uint32_t)s->N;
    f->g.d3 = s->d3 ? 1 : 0;
    f->g.n_types = (int)s->m.n_types;
    f->S = s->m.n_species;
    f->steps = s->traj_len;
    f->K = (uint32_t)std::min<uint64_t>(checkpoint_every, f->steps);
    const char *per_substep = std::getenv("GLAZIER_FIT_PER_SUBSTEP");
    f->fused = fit_fused_fits(f->g, f->max_sub) && !(per_substep && per_substep[0] == '1');
    const size_t fw = f->field_words();
    const uint64_t n_ckpt = (f->steps + f->K - 1) / f->K;
    FIT_TRY(hipStreamCreateWithFlags(&f->own, hipStreamNonBlocking));
    FIT_TRY(hipMalloc(&f->types, f->steps * s->N));
    FIT_TRY(hipMalloc(&f->ckpt, n_ckpt * fw * 4));
    FIT_TRY(hipMalloc(&f->seg, (size_t)f->K * fw * 4));
    if (!f->fused) FIT_TRY(hipMalloc(&f->sub, (size_t)f->max_sub * fw * 4));
    for (int i = 0; i < 2; ++i) {
        FIT_TRY(hipMalloc(&f->work[i], fw * 4));
        FIT_TRY(hipMalloc(&f->lam[i], fw * 4));
    }
    FIT_TRY(hipMalloc(&f->part, (size_t)f->S * f->part_blocks() * kPartWords * sizeof(double)));
    FIT_TRY(hipMalloc(&f->grad, (size_t)f->S * kPartWords * sizeof(double)));
    FIT_TRY(hipStreamSynchronize(s->stream));
    FIT_TRY(hipMemcpy(f->types, s->traj, f->steps * s->N, hipMemcpyDeviceToDevice));
#undef FIT_TRY
    *out = f;
    return GHIP_OK;
}

ghip_status ghip_fit_destroy(ghip_fit *f) {
    delete f;
    return GHIP_OK;
}

ghip_status ghip_fit_model_params(const ghip_fit *f, ghip_field_params *out) {
    if (!f || !out) return fail(GHIP_BAD_ARG, "null argument");
    std::memset(out, 0, sizeof *out);
    for (uint32_t k = 0; k < f->S; ++k) {
        out->diffusion[k] = f->m.species[k].diffusion;
        out->decay[k] = f->m.species[k].decay;
        for (uint32_t t = 0; t < f->m.n_types; ++t) {
            out->secretion[t][k] = f->m.secretion[t][k];
            out->uptake[t][k] = f->m.uptake[t][k];
        }
    }
    return GHIP_OK;
}

ghip_status ghip_fit_forward(ghip_fit *f, const ghip_field_params *p, const float *u0, float *u_final,
                             void *stream) {
    if (!f || !p || !u0 || !u_final) return fail(GHIP_BAD_ARG, "null argument");
    hipStream_t st = stream ? (hipStream_t)stream : f->own;
    f->coef = coef_of(f, *p);
    const size_t bytes = f->field_words() * 4;
    // checkpoint 0 is the initial field; the others are written as reached
    HIP_TRY(hipMemcpyAsync(f->ckpt, u0, bytes, hipMemcpyDeviceToDevice, st));
    const uint64_t n_seg = (f->steps + f->K - 1) / f->K;
    for (uint64_t sg = 0; sg < n_seg; ++sg) {
        const uint64_t m0 = sg * f->K;
        const uint64_t len = std::min<uint64_t>(f->K, f->steps - m0);
        float *last = m0 + len == f->steps ? u_final : f->ckpt + (sg + 1) * f->field_words();
        // the segment's inner states go to the segment buffer, free until a backward pass
        ghip_status s = forward_steps(f, f->ckpt + sg * f->field_words(), m0, len, f->seg + f->field_words(),
                                      last, st);
        if (s != GHIP_OK) return s;
    }
    f->have_forward = true;
    if (!stream) HIP_TRY(hipStreamSynchronize(st));
    return GHIP_OK;
}

}  // extern "C"

namespace {

// The backward pass with the partial sums in `part` (S * part_blocks *
// kPartWords doubles, cleared here) and their reduction in `red` (S *
// kPartWords); the host copy of the reduction stays on the stack.
ghip_status backward_into(ghip_fit *f, const float *grad_final, ghip_field_params *grad, float *grad_u0,
                          hipStream_t st, double *part, double *red) {
    const size_t fw = f->field_words();
    HIP_TRY(launch_fit_zero(part, (size_t)f->S * f->part_blocks() * kPartWords, st));
    HIP_TRY(hipMemcpyAsync(f->lam[0], grad_final, fw * 4, hipMemcpyDeviceToDevice, st));
    int li = 0;
    const uint64_t n_seg = (f->steps + f->K - 1) / f->K;
    for (uint64_t sg = n_seg; sg-- > 0;) {
        const uint64_t m0 = sg * f->K;
        const uint64_t len = std::min<uint64_t>(f->K, f->steps - m0);
        // the segment's step starts: the checkpoint, then states 1 .. len - 1
        const float *start0 = f->ckpt + sg * fw;
        if (len > 1) {
            ghip_status s = forward_steps(f, start0, m0, len - 1, f->seg + fw, f->seg + (len - 1) * fw, st);
            if (s != GHIP_OK) return s;
        }
        for (uint64_t j = len; j-- > 0;) {
            const uint64_t m = m0 + j;
            if (f->fused) {
                HIP_TRY(launch_fit_step_adjoint_fused(f->g, f->coef, f->S, f->max_sub, f->types + m * f->g.n,
                                                      j == 0 ? start0 : f->seg + j * fw, f->lam[li],
                                                      f->lam[li ^ 1], part, st));
                li ^= 1;
                continue;
            }
            // substep states: state s - 1 feeds substep s; state 0 is the step start
            const float *start = j == 0 ? start0 : f->seg + j * fw;
            for (int s = 1; s <= f->max_sub; ++s) {
                const float *from = s == 1 ? start : f->sub + (size_t)(s - 2) * fw;
                HIP_TRY(launch_fit_substep(f->g, f->coef, f->S, from, f->sub + (size_t)(s - 1) * fw, s, st));
            }
            const float *pre = f->sub + (size_t)(f->max_sub - 1) * fw;
            HIP_TRY(launch_fit_exchange_adjoint(f->g, f->coef, f->S, f->types + m * f->g.n, pre, f->lam[li],
                                                part, st));
            for (int s = f->max_sub; s >= 1; --s) {
                const float *before = s == 1 ? start : f->sub + (size_t)(s - 2) * fw;
                HIP_TRY(launch_fit_substep_adjoint(f->g, f->coef, f->S, f->lam[li], before, f->lam[li ^ 1],
                                                   part, s, st));
                li ^= 1;
            }
        }
    }
    HIP_TRY(hipMemcpyAsync(grad_u0, f->lam[li], fw * 4, hipMemcpyDeviceToDevice, st));
    HIP_TRY(launch_fit_reduce(f->S, f->part_blocks(), part, red, st));
    double acc[kMaxSpecies * kPartWords];
    HIP_TRY(hipMemcpyAsync(acc, red, (size_t)f->S * kPartWords * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::memset(grad, 0, sizeof *grad);
    for (uint32_t k = 0; k < f->S; ++k) {
        const double *w = acc + (size_t)k * kPartWords;
        const double sub = (double)f->m.species[k].substeps;
        grad->diffusion[k] = w[0] / sub;
        grad->decay[k] = w[1] / sub;
        for (uint32_t t = 0; t < f->m.n_types; ++t) {
            grad->secretion[t][k] = w[2 + 2 * t];
            grad->uptake[t][k] = w[3 + 2 * t];
        }
    }
    return GHIP_OK;
}

}  // namespace

extern "C" {

ghip_status ghip_fit_backward(ghip_fit *f, const float *grad_final, ghip_field_params *grad, float *grad_u0,
                              void *stream) {
    if (!f || !grad_final || !grad || !grad_u0) return fail(GHIP_BAD_ARG, "null argument");
    if (!f->have_forward) return fail(GHIP_BAD_ARG, "backward needs a forward pass first");
    hipStream_t st = stream ? (hipStream_t)stream : f->own;
    return backward_into(f, grad_final, grad, grad_u0, st, f->part, f->grad);
}

ghip_status ghip_fit_backward_noalloc(ghip_fit *f, const float *grad_final, ghip_field_params *grad,
                                      float *grad_u0, void *stream) {
    if (!f || !grad_final || !grad || !grad_u0) return fail(GHIP_BAD_ARG, "null argument");
    if (!f->have_forward) return fail(GHIP_BAD_ARG, "backward needs a forward pass first");
    const size_t part_words = (size_t)f->S * f->part_blocks() * kPartWords;
    const size_t need = part_words + (size_t)f->S * kPartWords;
    if (need > kFitScratchWords)
        return fail(GHIP_UNSUPPORTED, "this fit's backward pass needs " + std::to_string(need) +
                                          " words of scratch; the static scratch holds " +
                                          std::to_string(kFitScratchWords) + " (use ghip_fit_backward)");
    double *base = nullptr;
    HIP_TRY(fit_scratch(&base));
    hipStream_t st = stream ? (hipStream_t)stream : f->own;
    return backward_into(f, grad_final, grad, grad_u0, st, base, base + part_words);
}

}  // extern "C"
