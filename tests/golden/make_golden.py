"""Writes the committed fixtures of tests/golden (run from the repository root):

  chi2_49_lut256.npy    float32 scipy.stats.chi2.ppf(i/256, 49), i = 0..254: the inverse
                        chi-square LUT entries of src/compute/nlsar/stats.cpp:67-80, where the
                        reference calls GSL's gsl_cdf_chisq_Pinv (GSL is not available here;
                        scipy's ppf is an independent implementation of the same function).
  train32_stats.txt     NL-SAR statistics (patch 3,5,7 x scale 1,3) of the seeded homogeneous
                        32x32 training area used by the NL-SAR benchmark, in the library's
                        statistics text format, computed by the CPU oracle.
  example_crop.npz      rows/cols 0..127 of examples/insar_test_data.mat of the reference
                        (ampl_master, ampl_slave, dphase): a real interferogram as input data.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    from scipy.stats import chi2
    lut = np.array([chi2.ppf(i / 256.0, 49) for i in range(255)], dtype=np.float64).astype(np.float32)
    np.save(os.path.join(HERE, "chi2_49_lut256.npy"), lut)

    from oracle import oracle as O
    from despeckcl_amd import synthetic as S
    import despeckcl_amd as D
    a1, a2, ph = S.homogeneous_area(32, S.SEED_BASE + 100)
    st = O.nlsar_train_insar(a1, a2, ph, [3, 5, 7], [1, 3])
    coll = D.nlsar_stats_collection()
    for (p, s), v in st.items():
        x = D.stats(lut_size=v.lut_size)
        x.dissims_min, x.dissims_max = v.dissims_min, v.dissims_max
        x.quantilles, x.chi2cdf_inv = v.quantilles, v.chi2cdf_inv
        coll[D.params(p, s)] = x
    D.store_nlsar_stats_collection(coll, os.path.join(HERE, "train32_stats.txt"))

    mat = "/root/reference/examples/insar_test_data.mat"
    if os.path.exists(mat):
        import scipy.io as sio
        m = sio.loadmat(mat)
        np.savez_compressed(os.path.join(HERE, "example_crop.npz"),
                            ampl_master=np.ascontiguousarray(m["ampl_master"][:128, :128], np.float32),
                            ampl_slave=np.ascontiguousarray(m["ampl_slave"][:128, :128], np.float32),
                            dphase=np.ascontiguousarray(m["dphase"][:128, :128], np.float32))


if __name__ == "__main__":
    main()
