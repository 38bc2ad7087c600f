"""seispie_amd: the solver hot path of seispie (2-D staggered-grid FWI)
written for one AMD MI355X (gfx950).

Layout mirrors the reference package so that `[solver] method = fd2d`
style configs resolve the same way:
  seispie_amd.solver.fd2d     the solver plugin (seispie/solver/fd2d)
  seispie_amd.workflow        forward / adjoint / inversion drivers
  seispie_amd.optimize        L-BFGS / CG / GD with the reference line searches
  seispie_amd.autograd        torch.autograd.Function around forward + adjoint
  seispie_amd._lib            ctypes binding of lib/libseispie_hip.so
"""
__version__ = "0.1.0"
