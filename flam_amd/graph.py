"""rskelf_sv replayed from a HIP graph (rskx_sv_graph).

The solve of an FGMRES step is the same launch sequence every time; a
SolveGraph records it once for its own (nrhs, N) block and replays it:
solve_(Xt) copies Xt into the block, launches the graph and copies the
result back, in place like flam_amd.rskelf.rskelf_sv_ (one graph launch
instead of 2 + 2 nlvl kernel launches per solve)."""
import ctypes

import torch

from . import _lib as L
from .rskelf import FlamError, _check


class SolveGraph:
    def __init__(self, F, nrhs, dtype=torch.float64, trans="n"):
        if getattr(F, "cplx", False):
            raise FlamError(L.RSKX_ERR_UNSUP, "FLAM:rskx:graph: built for real factors")
        self.F = F
        self.nrhs = nrhs
        self.X = torch.zeros(nrhs, F.N, dtype=dtype, device="cuda")   # column-major N x nrhs
        dt = L.RSKX_F64 if dtype == torch.float64 else L.RSKX_F32
        _check(L.lib().rskx_set_stream(torch.cuda.current_stream().cuda_stream))
        h = ctypes.c_void_p()
        _check(L.lib().rskx_sv_graph(F.handle, self.X.data_ptr(), F.N, nrhs,
                                     str(trans).encode()[:1], dt, ctypes.byref(h)))
        self._h = h.value

    @property
    def cplx(self):
        return False

    def launch(self):
        """Solve in place on self.X (rows = right-hand sides)."""
        _check(L.lib().rskx_set_stream(torch.cuda.current_stream().cuda_stream))
        _check(L.lib().rskx_graph_launch(self._h))
        return self.X

    def solve_(self, Xt):
        """Xt (nrhs, N) <- F^{-1} Xt through the graph."""
        self.X.copy_(Xt)
        self.launch()
        Xt.copy_(self.X)
        return Xt

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            try:
                L.lib().rskx_graph_free(h)
            except Exception:
                pass
            self._h = None
