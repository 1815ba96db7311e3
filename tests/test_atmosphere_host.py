"""NSX atmosphere table loader (xpsi/Photosphere.py:167-218): a table written
in the documented file layout (rows over log T, log g, log10(E/kT), mu
descending; columns E, mu, log I, -, T, g) loads back into the 4-D table
layout (T, g, mu ascending, E) it was written from."""
import numpy as np

from xpsi_amd.atmosphere import load_NSX_table


def test_nsx_round_trip(tmp_path):
    rng = np.random.default_rng(8)
    logT = np.linspace(5.1, 6.8, 5)
    logg = np.linspace(13.7, 15.0, 3)
    mu = np.sort(rng.uniform(0.01, 1.0, 7))
    logE = np.linspace(-1.3, 2.0, 11)
    I = 10.0 ** rng.uniform(-3.0, 3.0, (5, 3, 7, 11))      # [T, g, mu, E]
    rows = []
    for i in range(5):
        for j in range(3):
            for k in range(11):
                for l in range(7):
                    m = 7 - 1 - l                           # mu descending within the file
                    rows.append([logE[k], mu[m], np.log10(I[i, j, m, k]), 0.0, logT[i], logg[j]])
    path = tmp_path / "nsx.out"
    np.savetxt(path, np.array(rows), fmt="%.17e")
    T, g, M, E, buf = load_NSX_table(str(path))
    np.testing.assert_array_equal(T, logT)
    np.testing.assert_array_equal(g, logg)
    np.testing.assert_array_equal(M, mu)
    np.testing.assert_array_equal(E, logE)
    np.testing.assert_allclose(buf, I.reshape(-1), rtol=1e-13)
    # other column choices
    sw = np.array(rows)[:, [4, 5, 2, 3, 0, 1]]
    np.savetxt(path, sw, fmt="%.17e")
    T2, g2, M2, E2, buf2 = load_NSX_table(str(path), Tcol=0, gcol=1, mucol=5, Ecol=4, spe_Icol=2)
    np.testing.assert_array_equal(M2, mu)
    np.testing.assert_allclose(buf2, buf, rtol=1e-15)
