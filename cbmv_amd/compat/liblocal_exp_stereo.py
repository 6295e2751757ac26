"""Stand-in for the reference's ``liblocal_exp_stereo`` (cpp/localexp/local_exp_stereo.cpp:486-491).

The graph-cut local-expansion optimiser stays outside this path; calling it
raises so that a caller cannot mistake it for a result.
"""


def run_local_exp_stereo(*args, **kwargs):
    raise NotImplementedError(
        "local expansion is not part of the MI355X pipeline; use the SGM/CBCA "
        "post-processing (isLocalExp=False)")
