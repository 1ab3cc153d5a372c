"""Selective Adam: the fused per-row Adam step of the reference
(``textured_gaussians/optimizers/selective_adam.py:48-88``, kernel
``cuda/csrc/adam.cu``), running the gfx950 kernel.  No bias correction, as in
the reference."""
import torch

from .cuda._wrapper import selective_adam_update


class SelectiveAdam(torch.optim.Adam):
    """Adam that updates only the rows (first dimension) whose ``visibility``
    flag is set; each param group must hold exactly one tensor."""

    def __init__(self, params, eps, betas):
        super().__init__(params=params, eps=eps, betas=betas)

    @torch.no_grad()
    def step(self, visibility):
        N = visibility.numel()
        for group in self.param_groups:
            lr = group["lr"]
            eps = group["eps"]
            beta1, beta2 = group["betas"]
            assert len(group["params"]) == 1, "more than one tensor in group"
            param = group["params"][0]
            if param.grad is None:
                continue
            state = self.state[param]
            if len(state) == 0:
                state["step"] = torch.tensor(0.0, dtype=torch.float32)
                state["exp_avg"] = torch.zeros_like(param, memory_format=torch.preserve_format)
                state["exp_avg_sq"] = torch.zeros_like(param, memory_format=torch.preserve_format)
            M = param.numel() // N
            selective_adam_update(param, param.grad, state["exp_avg"], state["exp_avg_sq"], visibility, lr, beta1,
                                  beta2, eps, N, M)
