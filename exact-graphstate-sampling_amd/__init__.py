"""MI355X-native engine for the in-place linear-rank-width dynamic program of exact graph-state sampling.

Host side (pure Python, bit-exact planning): ``tableau``, ``graphform``, ``amp``, ``plan``, ``inplace``,
``engine`` (plan/job writers), ``planio`` (file formats) and ``instances`` (banded random instances).
Device side: ``gsamp`` loads the HIP engine (``libgsamp_amd.so``, C ABI declared in include/gsamp_amd.h) and
exposes the amplitude, marginal and sampling entry points, the batched scorer and its adjoint; ``adjoint`` maps the
adjoint onto rotation angles (and a torch.autograd.Function); ``autotune`` picks the batched engine's tiles and
schedules.
"""
