"""
Compute-engine interface: :class:`KernelModel` and :class:`Kernel`.

Same surface and normalisation as the reference plugin boundary
(sasmodels/kernel.py:13-205):

* ``KernelModel.make_kernel(q_vectors)`` returns a :class:`Kernel` bound to
  the q points (``[q]`` for 1D, ``[qx, qy]`` for 2D);
* ``Kernel.Fq`` calls ``_call_kernel`` and turns the raw accumulators in
  ``self.result`` into ``(F1, F2, R_eff, V_shell, V_form/V_shell)``:
  every sum is divided by the total weight (treated as 1 when it is 0) and
  the shell volume is treated as 1 when it is 0 (kernel.py:170-188);
* ``Kernel.Iq = scale/V_shell * <F^2> + background`` (kernel.py:80-99).
"""


class KernelModel:
    """Model definition for a compute engine."""
    info = None
    dtype = None

    def make_kernel(self, q_vectors):
        """Instantiate a kernel evaluating the model at *q_vectors*."""
        raise NotImplementedError("need to implement make_kernel")

    def release(self):
        """Free resources associated with the model."""
        pass


class Kernel:
    """
    A model bound to q points.  Subclasses set ``q_input`` (with ``nq``),
    ``info``, ``dim`` ('1d' or '2d'), ``dtype`` and ``result``, and implement
    ``_call_kernel``.
    """
    dim = None
    info = None
    dtype = None
    q_input = None
    result = None

    def Iq(self, call_details, values, cutoff, magnetic):
        """I(q) = scale * <F^2>/<V> + background."""
        _, F2, _, shell_volume, _ = self.Fq(call_details, values, cutoff,
                                            magnetic, radius_effective_mode=0)
        combined_scale = values[0]/shell_volume
        background = values[1]
        return combined_scale*F2 + background
    __call__ = Iq

    def Fq(self, call_details, values, cutoff, magnetic, radius_effective_mode=0):
        """
        Returns <F(q)> (None for 2D), <F(q)^2>, effective radius, shell
        volume and form:shell volume ratio, all averaged over the mesh.
        """
        self._call_kernel(call_details, values, cutoff, magnetic,
                          radius_effective_mode)
        nq = self.q_input.nq
        nout = 2 if self.info.have_Fq and self.dim == "1d" else 1
        total_weight = self.result[nout*nq + 0]
        if total_weight == 0.:
            total_weight = 1.
        form_volume = self.result[nout*nq + 1]/total_weight
        shell_volume = self.result[nout*nq + 2]/total_weight
        radius_effective = self.result[nout*nq + 3]/total_weight
        if shell_volume == 0.:
            shell_volume = 1.
        F1 = (self.result[1:nout*nq:nout]/total_weight if nout == 2 else None)
        F2 = self.result[0:nout*nq:nout]/total_weight
        return F1, F2, radius_effective, shell_volume, form_volume/shell_volume

    def release(self):
        """Free resources associated with the kernel instance."""
        pass

    def _call_kernel(self, call_details, values, cutoff, magnetic, radius_effective_mode):
        """Fill ``self.result`` with the raw accumulators."""
        raise NotImplementedError()
