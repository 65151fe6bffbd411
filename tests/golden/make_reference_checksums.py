"""Extract the SHA-256 checksum fixtures of the reference's own tests into reference_checksums.json.

Source: tests/reference_outputs/out_*.json of k-wave-python (HDF5 input-file
summaries written by its legacy C++-input path; recipe tests/h5_summary.py:26-33:
float32 array in HDF5 (MATLAB F-order) layout, ``round(8) + 0``, sha256 of the
buffer).  Only the datasets that pin this package's host set-up are kept: the
smoothed initial pressure (make_disc / make_ball + smooth), c0, rho0 and the
staggered densities (the ``_stagger`` midpoint rule); and every variable of four
example input files, which pin the C++-binary file route.  Data only: no
reference source is copied.  Run from the repository root with the reference checked out
at /root/reference.
"""
import json
import os
import sys

SRC = os.environ.get("KWAVE_REFERENCE_OUTPUTS", "/root/reference/tests/reference_outputs")
KEEP = {
    "out_ivp_homogeneous_medium": ["p0_source_input", "c0", "rho0", "rho0_sgx", "rho0_sgy"],
    "out_ivp_heterogeneous_medium": ["p0_source_input", "c0", "rho0", "rho0_sgx", "rho0_sgy"],
    "out_ivp_3D_simulation": ["p0_source_input", "c0", "rho0", "rho0_sgx", "rho0_sgy", "rho0_sgz"],
}


# every variable of these example input files: pins the C++-binary file route
# (kwave_amd.solvers.cpp_simulation.input_variables)
CPP_FILES = ["out_ivp_homogeneous_medium", "out_ivp_heterogeneous_medium", "out_ivp_3D_simulation",
             "out_tvsp_homogeneous_medium_monopole"]


def main(out=os.path.join(os.path.dirname(__file__), "reference_checksums.json")):
    res = {}
    for name, keys in KEEP.items():
        with open(os.path.join(SRC, name + ".json")) as f:
            d = json.load(f)
        res[name] = {k: {"dtype": d[k]["dtype"], "shape": d[k]["shape"], "sha256_round8": d[k]["checksums"]["8"]} for k in keys}
    res["cpp_input_files"] = {}
    for name in CPP_FILES:
        with open(os.path.join(SRC, name + ".json")) as f:
            d = json.load(f)
        res["cpp_input_files"][name] = {k: {"dtype": v["dtype"], "shape": v["shape"], "sha256_round8": v["checksums"]["8"]}
                                        for k, v in d.items() if k != "root"}
    with open(out, "w") as f:
        json.dump(res, f, indent=1)
    print(f"wrote {out}", file=sys.stderr)


if __name__ == "__main__":
    main()
