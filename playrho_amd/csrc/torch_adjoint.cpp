// libtorch binding of the adjoint (prx_adjoint_*): a
// torch::autograd::Function registered as the operator
// torch.ops.playrho_amd.billiards_loss.  The tensors carry parameters and
// gradients in and out; the computation is the library's (forward steps with
// checkpoints, recorded recomputation and the reversed steps, all on the
// device).
//
//   billiards_loss(v0, friction, restitution, descs, desc_offsets, targets,
//                  steps, every) -> loss
//
//   v0           (W, B, 3)  initial velocity (vx, vy, w) of every body of
//                           world w (static bodies' entries are ignored, as
//                           SetVelocity ignores them, WorldBody.cpp)
//   friction     (W, S)     per shape (SetFriction, WorldShape.cpp:79-84)
//   restitution  (W, S)     per shape (SetRestitution, WorldShape.cpp:91-96)
//   descs        uint8      the W world descriptions, concatenated
//   desc_offsets int64      (W + 1) byte offsets of the descriptions
//   targets      (W, B, 2)  target positions
//   loss         (W,)       sum over dynamic bodies of |x_K - target|^2
//
// Gradients flow to v0, friction and restitution; every world must have B
// bodies and S shapes.
//
//   articulated_loss(v0, friction, restitution, torque, descs, desc_offsets,
//                    targets, steps, every) -> loss
//
// is the same with torque (W, J): every joint's maximum motor torque, also
// differentiated.
//
//   simulate(v0, friction, restitution, torque, descs, desc_offsets, steps,
//            every) -> state
//
// returns the final state (W, B, 6): x, y, angle, vx, vy, w of every body
// after `steps` steps (torque may be (W, 0) for worlds without joints).  Its
// backward takes any upstream gradient dL/dstate: it steps the worlds again
// from their checkpoints and reverses the steps with that gradient as the
// seed, so the loss is the caller's; a forward-only call pays for the forward
// pass alone.
#include <torch/library.h>
#include <torch/torch.h>

#include <cstdint>
#include <cstring>
#include <vector>

#include "playrho_amd.h"

namespace {

using torch::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

void Check(int status, const char* what) {
    TORCH_CHECK(status == PRX_OK, "playrho_amd: ", what, ": ", prx_last_error());
}

// Owns the worlds, the batch and the adjoint of one call.
struct Run {
    std::vector<prx_world*> worlds;
    prx_batch* batch = nullptr;
    prx_adjoint* adj = nullptr;
    ~Run() {
        if (adj) prx_adjoint_destroy(adj);
        if (batch) prx_batch_destroy(batch);
        for (prx_world* w : worlds) prx_world_destroy(w);
    }
};

class BilliardsLoss : public torch::autograd::Function<BilliardsLoss> {
public:
    static Tensor forward(AutogradContext* ctx, const Tensor& v0, const Tensor& friction, const Tensor& restitution,
                          const Tensor& descs, const Tensor& desc_offsets, const Tensor& targets, int64_t steps,
                          int64_t every) {
        TORCH_CHECK(v0.dim() == 3 && v0.size(2) == 3, "v0 must be (worlds, bodies, 3)");
        const int64_t W = v0.size(0), B = v0.size(1);
        TORCH_CHECK(friction.dim() == 2 && friction.size(0) == W, "friction must be (worlds, shapes)");
        TORCH_CHECK(restitution.sizes() == friction.sizes(), "restitution must match friction's shape");
        const int64_t S = friction.size(1);
        TORCH_CHECK(targets.numel() == W * B * 2, "targets must hold (worlds, bodies, 2) values");
        TORCH_CHECK(descs.scalar_type() == torch::kUInt8, "descs must be uint8");
        TORCH_CHECK(desc_offsets.numel() == W + 1, "desc_offsets must hold worlds + 1 offsets");
        TORCH_CHECK(steps > 0 && every > 0, "steps and every must be positive");

        const auto cpu = torch::TensorOptions().dtype(torch::kFloat32).device(torch::kCPU);
        const Tensor v = v0.detach().to(cpu).contiguous();
        const Tensor f = friction.detach().to(cpu).contiguous();
        const Tensor e = restitution.detach().to(cpu).contiguous();
        const Tensor tg = targets.detach().to(cpu).contiguous();
        const Tensor d = descs.detach().to(torch::kCPU).contiguous();
        const Tensor off = desc_offsets.detach().to(torch::kCPU, torch::kInt64).contiguous();
        const float* pv = v.data_ptr<float>();
        const float* pf = f.data_ptr<float>();
        const float* pe = e.data_ptr<float>();
        const std::uint8_t* pd = d.data_ptr<std::uint8_t>();
        const int64_t* po = off.data_ptr<int64_t>();

        Run run;
        for (int64_t w = 0; w < W; ++w) {
            prx_world* h = nullptr;
            Check(prx_world_create(nullptr, &h), "world create");
            run.worlds.push_back(h);
            TORCH_CHECK(po[w] >= 0 && po[w] <= po[w + 1] && po[w + 1] <= d.numel(), "bad description offsets");
            Check(prx_world_load_desc(h, pd + po[w], std::size_t(po[w + 1] - po[w])), "load description");
            TORCH_CHECK(int64_t(prx_world_body_count(h)) == B, "world ", w, " has ", prx_world_body_count(h),
                        " bodies, v0 has ", B);
            TORCH_CHECK(int64_t(prx_world_shape_count(h)) == S, "world ", w, " has ", prx_world_shape_count(h),
                        " shapes, friction has ", S);
            for (int64_t b = 0; b < B; ++b)
                Check(prx_world_set_velocity(h, std::uint32_t(b), pv + (w * B + b) * 3), "set velocity");
            for (int64_t s = 0; s < S; ++s) {
                Check(prx_shape_set_friction(h, std::uint32_t(s), pf[w * S + s]), "set friction");
                Check(prx_shape_set_restitution(h, std::uint32_t(s), pe[w * S + s]), "set restitution");
            }
        }
        prx_batch_conf bc;
        prx_batch_conf_default(&bc);
        bc.use_graph = 0;
        Check(prx_batch_create(run.worlds.data(), int(W), &bc, &run.batch), "batch create");
        prx_step_conf sc;
        prx_step_conf_default(&sc);
        sc.do_toi = 0;  // the adjoint differentiates the regular step
        Check(prx_adjoint_create(run.batch, int(steps), int(every), &sc, &run.adj), "adjoint create");

        Tensor loss = torch::empty({W}, cpu);
        Tensor gv = torch::empty({W, B, 3}, cpu);
        Tensor gf = torch::empty({W, S}, cpu);
        Tensor ge = torch::empty({W, S}, cpu);
        Check(prx_adjoint_run(run.adj, &sc, tg.data_ptr<float>(), loss.data_ptr<float>(), gv.data_ptr<float>(),
                              gf.data_ptr<float>(), ge.data_ptr<float>()),
              "adjoint run");
        ctx->save_for_backward({gv.to(v0.options()), gf.to(friction.options()), ge.to(restitution.options())});
        return loss.to(v0.options());
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grad_out) {
        const auto saved = ctx->get_saved_variables();
        const Tensor g = grad_out[0];
        return {saved[0] * g.to(saved[0].options()).view({-1, 1, 1}),
                saved[1] * g.to(saved[1].options()).view({-1, 1}),
                saved[2] * g.to(saved[2].options()).view({-1, 1}),
                Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

Tensor billiards_loss(const Tensor& v0, const Tensor& friction, const Tensor& restitution, const Tensor& descs,
                      const Tensor& desc_offsets, const Tensor& targets, int64_t steps, int64_t every) {
    return BilliardsLoss::apply(v0, friction, restitution, descs, desc_offsets, targets, steps, every);
}

// Byte offsets of the joint records' maximum motor torque in a world
// description (layout of prx_world_save_desc: shapes, bodies, then joints of
// 60 bytes each: bodies and flags, 12 floats with the torque the ninth).
std::vector<std::size_t> TorqueOffsets(const std::uint8_t* d, std::size_t len) {
    auto u32 = [&](std::size_t o) {
        TORCH_CHECK(o + 4 <= len, "truncated world description");
        std::uint32_t v;
        std::memcpy(&v, d + o, 4);
        return v;
    };
    std::size_t off = 4;
    const std::uint32_t ns = u32(0);
    for (std::uint32_t i = 0; i < ns; ++i) off += 4 + 16 + 12 + 16 * std::size_t(u32(off));
    const std::uint32_t nb = u32(off);
    off += 4;
    for (std::uint32_t i = 0; i < nb; ++i) off += 84 + 4 + 4 * std::size_t(u32(off + 84));
    const std::uint32_t nj = u32(off);
    off += 4;
    std::vector<std::size_t> out;
    for (std::uint32_t k = 0; k < nj; ++k) out.push_back(off + 60 * std::size_t(k) + 44);
    TORCH_CHECK(out.empty() || out.back() + 4 <= len, "truncated world description");
    return out;
}

// Articulated worlds: the same loss, also differentiable with respect to
// every joint's maximum motor torque (RevoluteJointConf::maxMotorTorque).
// Each world is built as CreateJoint with that torque would build it.
class ArticulatedLoss : public torch::autograd::Function<ArticulatedLoss> {
public:
    static Tensor forward(AutogradContext* ctx, const Tensor& v0, const Tensor& friction, const Tensor& restitution,
                          const Tensor& torque, const Tensor& descs, const Tensor& desc_offsets,
                          const Tensor& targets, int64_t steps, int64_t every) {
        TORCH_CHECK(v0.dim() == 3 && v0.size(2) == 3, "v0 must be (worlds, bodies, 3)");
        const int64_t W = v0.size(0), B = v0.size(1);
        TORCH_CHECK(friction.dim() == 2 && friction.size(0) == W, "friction must be (worlds, shapes)");
        TORCH_CHECK(restitution.sizes() == friction.sizes(), "restitution must match friction's shape");
        TORCH_CHECK(torque.dim() == 2 && torque.size(0) == W, "torque must be (worlds, joints)");
        const int64_t S = friction.size(1), J = torque.size(1);
        TORCH_CHECK(targets.numel() == W * B * 2, "targets must hold (worlds, bodies, 2) values");
        TORCH_CHECK(descs.scalar_type() == torch::kUInt8, "descs must be uint8");
        TORCH_CHECK(desc_offsets.numel() == W + 1, "desc_offsets must hold worlds + 1 offsets");
        TORCH_CHECK(steps > 0 && every > 0, "steps and every must be positive");
        const auto cpu = torch::TensorOptions().dtype(torch::kFloat32).device(torch::kCPU);
        const Tensor v = v0.detach().to(cpu).contiguous();
        const Tensor f = friction.detach().to(cpu).contiguous();
        const Tensor e = restitution.detach().to(cpu).contiguous();
        const Tensor tq = torque.detach().to(cpu).contiguous();
        const Tensor tg = targets.detach().to(cpu).contiguous();
        const Tensor d = descs.detach().to(torch::kCPU).contiguous();
        const Tensor off = desc_offsets.detach().to(torch::kCPU, torch::kInt64).contiguous();
        const float* pv = v.data_ptr<float>();
        const float* pf = f.data_ptr<float>();
        const float* pe = e.data_ptr<float>();
        const float* pt = tq.data_ptr<float>();
        const std::uint8_t* pd = d.data_ptr<std::uint8_t>();
        const int64_t* po = off.data_ptr<int64_t>();

        Run run;
        for (int64_t w = 0; w < W; ++w) {
            TORCH_CHECK(po[w] >= 0 && po[w] <= po[w + 1] && po[w + 1] <= d.numel(), "bad description offsets");
            std::vector<std::uint8_t> desc(pd + po[w], pd + po[w + 1]);
            const auto toff = TorqueOffsets(desc.data(), desc.size());
            TORCH_CHECK(int64_t(toff.size()) == J, "world ", w, " has ", toff.size(), " joints, torque has ", J);
            for (int64_t j = 0; j < J; ++j) std::memcpy(desc.data() + toff[std::size_t(j)], pt + w * J + j, 4);
            prx_world* h = nullptr;
            Check(prx_world_create(nullptr, &h), "world create");
            run.worlds.push_back(h);
            Check(prx_world_load_desc(h, desc.data(), desc.size()), "load description");
            TORCH_CHECK(int64_t(prx_world_body_count(h)) == B, "world ", w, " body count differs from v0");
            TORCH_CHECK(int64_t(prx_world_shape_count(h)) == S, "world ", w, " shape count differs from friction");
            for (int64_t b = 0; b < B; ++b)
                Check(prx_world_set_velocity(h, std::uint32_t(b), pv + (w * B + b) * 3), "set velocity");
            for (int64_t s2 = 0; s2 < S; ++s2) {
                Check(prx_shape_set_friction(h, std::uint32_t(s2), pf[w * S + s2]), "set friction");
                Check(prx_shape_set_restitution(h, std::uint32_t(s2), pe[w * S + s2]), "set restitution");
            }
        }
        prx_batch_conf bc;
        prx_batch_conf_default(&bc);
        bc.use_graph = 0;
        Check(prx_batch_create(run.worlds.data(), int(W), &bc, &run.batch), "batch create");
        prx_step_conf sc;
        prx_step_conf_default(&sc);
        sc.do_toi = 0;  // the adjoint differentiates the regular step
        Check(prx_adjoint_create(run.batch, int(steps), int(every), &sc, &run.adj), "adjoint create");
        Tensor loss = torch::empty({W}, cpu);
        Tensor gv = torch::empty({W, B, 3}, cpu);
        Tensor gf = torch::empty({W, S}, cpu);
        Tensor ge = torch::empty({W, S}, cpu);
        Tensor gt = torch::zeros({W, J}, cpu);
        Check(prx_adjoint_run(run.adj, &sc, tg.data_ptr<float>(), loss.data_ptr<float>(), gv.data_ptr<float>(),
                              gf.data_ptr<float>(), ge.data_ptr<float>()),
              "adjoint run");
        if (J) Check(prx_adjoint_joint_gradients(run.adj, gt.data_ptr<float>()), "joint gradients");
        ctx->save_for_backward({gv.to(v0.options()), gf.to(friction.options()), ge.to(restitution.options()),
                                gt.to(torque.options())});
        return loss.to(v0.options());
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grad_out) {
        const auto saved = ctx->get_saved_variables();
        const Tensor g = grad_out[0];
        return {saved[0] * g.to(saved[0].options()).view({-1, 1, 1}),
                saved[1] * g.to(saved[1].options()).view({-1, 1}),
                saved[2] * g.to(saved[2].options()).view({-1, 1}),
                saved[3] * g.to(saved[3].options()).view({-1, 1}),
                Tensor(), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

Tensor articulated_loss(const Tensor& v0, const Tensor& friction, const Tensor& restitution, const Tensor& torque,
                        const Tensor& descs, const Tensor& desc_offsets, const Tensor& targets, int64_t steps,
                        int64_t every) {
    return ArticulatedLoss::apply(v0, friction, restitution, torque, descs, desc_offsets, targets, steps, every);
}

// The worlds, batch and adjoint of a simulate call, built from its inputs
// (each world as its description with the given torques, velocities and
// materials).
void BuildRun(Run& run, const Tensor& v, const Tensor& f, const Tensor& e, const Tensor& tq, const Tensor& d,
              const Tensor& off, int64_t W, int64_t B, int64_t S, int64_t J, int64_t steps, int64_t every,
              prx_step_conf& sc) {
    const float* pv = v.data_ptr<float>();
    const float* pf = f.data_ptr<float>();
    const float* pe = e.data_ptr<float>();
    const float* pt = J ? tq.data_ptr<float>() : nullptr;
    const std::uint8_t* pd = d.data_ptr<std::uint8_t>();
    const int64_t* po = off.data_ptr<int64_t>();
    for (int64_t w = 0; w < W; ++w) {
        TORCH_CHECK(po[w] >= 0 && po[w] <= po[w + 1] && po[w + 1] <= d.numel(), "bad description offsets");
        std::vector<std::uint8_t> desc(pd + po[w], pd + po[w + 1]);
        if (J) {
            const auto toff = TorqueOffsets(desc.data(), desc.size());
            TORCH_CHECK(int64_t(toff.size()) == J, "world ", w, " has ", toff.size(), " joints, torque has ", J);
            for (int64_t j = 0; j < J; ++j) std::memcpy(desc.data() + toff[std::size_t(j)], pt + w * J + j, 4);
        }
        prx_world* h = nullptr;
        Check(prx_world_create(nullptr, &h), "world create");
        run.worlds.push_back(h);
        Check(prx_world_load_desc(h, desc.data(), desc.size()), "load description");
        TORCH_CHECK(int64_t(prx_world_body_count(h)) == B, "world ", w, " body count differs from v0");
        TORCH_CHECK(int64_t(prx_world_shape_count(h)) == S, "world ", w, " shape count differs from friction");
        for (int64_t b = 0; b < B; ++b)
            Check(prx_world_set_velocity(h, std::uint32_t(b), pv + (w * B + b) * 3), "set velocity");
        for (int64_t s2 = 0; s2 < S; ++s2) {
            Check(prx_shape_set_friction(h, std::uint32_t(s2), pf[w * S + s2]), "set friction");
            Check(prx_shape_set_restitution(h, std::uint32_t(s2), pe[w * S + s2]), "set restitution");
        }
    }
    prx_batch_conf bc;
    prx_batch_conf_default(&bc);
    bc.use_graph = 0;
    Check(prx_batch_create(run.worlds.data(), int(W), &bc, &run.batch), "batch create");
    prx_step_conf_default(&sc);
    sc.do_toi = 0;  // the adjoint differentiates the regular step
    Check(prx_adjoint_create(run.batch, int(steps), int(every), &sc, &run.adj), "adjoint create");
}

class Simulate : public torch::autograd::Function<Simulate> {
public:
    static Tensor forward(AutogradContext* ctx, const Tensor& v0, const Tensor& friction, const Tensor& restitution,
                          const Tensor& torque, const Tensor& descs, const Tensor& desc_offsets, int64_t steps,
                          int64_t every) {
        TORCH_CHECK(v0.dim() == 3 && v0.size(2) == 3, "v0 must be (worlds, bodies, 3)");
        const int64_t W = v0.size(0), B = v0.size(1);
        TORCH_CHECK(friction.dim() == 2 && friction.size(0) == W, "friction must be (worlds, shapes)");
        TORCH_CHECK(restitution.sizes() == friction.sizes(), "restitution must match friction's shape");
        TORCH_CHECK(torque.dim() == 2 && torque.size(0) == W, "torque must be (worlds, joints)");
        TORCH_CHECK(descs.scalar_type() == torch::kUInt8, "descs must be uint8");
        TORCH_CHECK(desc_offsets.numel() == W + 1, "desc_offsets must hold worlds + 1 offsets");
        TORCH_CHECK(steps > 0 && every > 0, "steps and every must be positive");
        const auto cpu = torch::TensorOptions().dtype(torch::kFloat32).device(torch::kCPU);
        const Tensor v = v0.detach().to(cpu).contiguous();
        const Tensor f = friction.detach().to(cpu).contiguous();
        const Tensor e = restitution.detach().to(cpu).contiguous();
        const Tensor tq = torque.detach().to(cpu).contiguous();
        const Tensor d = descs.detach().to(torch::kCPU).contiguous();
        const Tensor off = desc_offsets.detach().to(torch::kCPU, torch::kInt64).contiguous();
        Run run;
        prx_step_conf sc;
        BuildRun(run, v, f, e, tq, d, off, W, B, friction.size(1), torque.size(1), steps, every, sc);
        Tensor state = torch::empty({W, B, 6}, cpu);
        Check(prx_adjoint_forward(run.adj, &sc, state.data_ptr<float>()), "adjoint forward");
        ctx->save_for_backward({v, f, e, tq, d, off});
        ctx->saved_data["steps"] = steps;
        ctx->saved_data["every"] = every;
        return state.to(v0.options());
    }

    static tensor_list backward(AutogradContext* ctx, tensor_list grad_out) {
        const auto sv = ctx->get_saved_variables();
        const Tensor &v = sv[0], &f = sv[1], &e = sv[2], &tq = sv[3], &d = sv[4], &off = sv[5];
        const int64_t W = v.size(0), B = v.size(1), S = f.size(1), J = tq.size(1);
        const int64_t steps = ctx->saved_data["steps"].toInt(), every = ctx->saved_data["every"].toInt();
        const auto cpu = torch::TensorOptions().dtype(torch::kFloat32).device(torch::kCPU);
        const Tensor sbar = grad_out[0].detach().to(cpu).contiguous();
        TORCH_CHECK(sbar.numel() == W * B * 6, "the state gradient must be (worlds, bodies, 6)");
        Run run;
        prx_step_conf sc;
        BuildRun(run, v, f, e, tq, d, off, W, B, S, J, steps, every, sc);
        Check(prx_adjoint_forward(run.adj, &sc, nullptr), "adjoint forward");
        Tensor gv = torch::empty({W, B, 3}, cpu);
        Tensor gf = torch::empty({W, S}, cpu);
        Tensor ge = torch::empty({W, S}, cpu);
        Tensor gt = torch::zeros({W, J}, cpu);
        Check(prx_adjoint_backward(run.adj, &sc, sbar.data_ptr<float>(), gv.data_ptr<float>(), gf.data_ptr<float>(),
                                   ge.data_ptr<float>()),
              "adjoint backward");
        if (J) Check(prx_adjoint_joint_gradients(run.adj, gt.data_ptr<float>()), "joint gradients");
        const auto like = grad_out[0].options();
        return {gv.to(like), gf.to(like), ge.to(like), gt.to(like), Tensor(), Tensor(), Tensor(), Tensor()};
    }
};

Tensor simulate(const Tensor& v0, const Tensor& friction, const Tensor& restitution, const Tensor& torque,
                const Tensor& descs, const Tensor& desc_offsets, int64_t steps, int64_t every) {
    return Simulate::apply(v0, friction, restitution, torque, descs, desc_offsets, steps, every);
}

}  // namespace

TORCH_LIBRARY(playrho_amd, m) {
    m.def("billiards_loss(Tensor v0, Tensor friction, Tensor restitution, Tensor descs, Tensor desc_offsets, "
          "Tensor targets, int steps, int every) -> Tensor");
    m.def("articulated_loss(Tensor v0, Tensor friction, Tensor restitution, Tensor torque, Tensor descs, "
          "Tensor desc_offsets, Tensor targets, int steps, int every) -> Tensor");
    m.def("simulate(Tensor v0, Tensor friction, Tensor restitution, Tensor torque, Tensor descs, "
          "Tensor desc_offsets, int steps, int every) -> Tensor");
}

TORCH_LIBRARY_IMPL(playrho_amd, CompositeImplicitAutograd, m) {
    m.impl("billiards_loss", billiards_loss);
    m.impl("articulated_loss", articulated_loss);
    m.impl("simulate", simulate);
}
