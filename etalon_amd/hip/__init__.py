"""The HIP side: ctypes binding (``_lib``), device arrays and the protocol backend
(``backend``), and fused time-stepping plans (``plan``)."""
