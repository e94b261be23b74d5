"""CPU oracle for the HDG Newton-GMRES path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, and only as the
checker (or the timed CPU baseline), never as the thing that is measured or
shipped.  The product path (``exasim_amd`` + ``libexahip.so``) never calls it.

``hdg`` is a numpy restatement of the reference's algorithm; every function
cites the reference file:line it follows.  ``models`` restates the pointwise
model functions that text2code would generate for the pdemodel files the
configurations use.
"""
