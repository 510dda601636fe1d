"""Drop-in boundary of the C++ mirror (host/alto_amd.hpp), on the CPU: a
host program written against the reference's own signatures -- the calls
of src/main.cpp:335-398 (create_alto, cpd_alto, cpstream with a
cpd_constraint and the dynamic-batching arguments), mttkrp_alto_par with
its omp_lock_t* / ofibs / oidx arguments (alto.hpp:904), create_da_mem /
destroy_da_mem (alto.hpp:606-679), init_salto / update_salto
(alto.hpp:151-299), cpd_fit_alto (cpd.hpp:193-275), cpd_error
(streaming_cpd.hpp:88) and the AltoTensor cr_masks / alto_cr_mask fields
(alto.hpp:53-54) -- compiles and links unchanged against libalto_amd.so, for
64- and 128-bit keys (ALTO_MASK_LENGTH) and with GENERATE_CR_BITS.  Without a
GPU the program's first device call must fail loudly (no CPU fallback)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "alto-stream_amd")
LIBDIR = os.path.join(PKG, "_build")

PROGRAM = r'''
#include "alto_amd.hpp"
#include <stdio.h>
#include <vector>
int main(int argc, char **argv) {
    SparseTensor *X;
    ImportSparseTensor(argv[1], TEXT_FORMAT, &X, 0);
    const IType rank = 4;
    KruskalModel *M;
    CreateKruskalModel(X->nmodes, X->dims, rank, &M);
    KruskalModelRandomInit(M, 1);
    Model model = argc > 2 ? SPCPSTREAM_ALTO : ALS;

    AltoTensor<LIType> *AT;
    create_alto(X, &AT, 4);
    printf("cr modes:");
    for (int n = 0; n < AT->nmode; ++n) printf(" %d", AT->cr_masks[n] != 0);
    printf(" union %d\n", AT->alto_cr_mask != 0);

    FType **ofibs = NULL;
    create_da_mem(-1, rank, AT, &ofibs);
    omp_lock_t *wlocks = NULL;
    std::vector<std::vector<size_t>> oidx;
    for (int n = 0; n < X->nmodes; ++n) mttkrp_alto_par(n, M->U, rank, AT, wlocks, ofibs, oidx);
    destroy_da_mem(AT, ofibs, rank, -1);
    cpd_alto(AT, M, 2, 1e-5);

    Matrix **grams = (Matrix **)malloc(sizeof(Matrix *) * X->nmodes);
    for (int n = 0; n < X->nmodes; ++n) {
        grams[n] = (Matrix *)malloc(sizeof(Matrix));
        grams[n]->I = grams[n]->J = rank;
        grams[n]->vals = (FType *)calloc(rank * rank, sizeof(FType));
        grams[n]->rowmajor = 1;
    }
    FType *mt = (FType *)calloc(X->dims[X->nmodes - 1] * rank, sizeof(FType));
    printf("fit %f\n", cpd_fit_alto(AT, M, grams, mt, 1.0));
    destroy_alto(AT);

    AltoTensor<LIType> *ST;
    init_salto(X, &ST, 2, 0);
    update_salto(X, ST, 2);
    destroy_alto(ST);
    printf("err %f\n", cpd_error(X, M));

    cpd_constraint *con = init_constraint();
    apply_nonneg_constraint(con);
    cpstream(X, (int)rank, 2, 0, 1e-5, 1, con, model == SPCPSTREAM_ALTO, 100000, 2);
    free_constraint(con);
    DestroyKruskalModel(M);
    return 0;
}
'''


def _build(tmp_path, defines):
    if not os.path.exists(os.path.join(LIBDIR, "libalto_amd.so")):
        pytest.skip("libalto_amd.so is not built")
    src = tmp_path / "main.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "main"
    subprocess.run(["g++", "-std=c++17", "-O1", "-fopenmp", *defines, "-I", os.path.join(PKG, "host"),
                    "-I", os.path.join(ROOT, "include"), str(src), os.path.join(PKG, "host", "alto_amd.cpp"),
                    "-L", LIBDIR, "-lalto_amd", "-Wl,-rpath," + LIBDIR, "-o", str(exe)], check=True)
    return str(exe)


DEFINES = [[], ["-DALTO_MASK_LENGTH=128"], ["-DGENERATE_CR_BITS"], ["-DALTO_MASK_LENGTH=128", "-DGENERATE_CR_BITS"]]


@pytest.mark.parametrize("defines", DEFINES)
def test_reference_style_program_builds_and_fails_loudly_without_gpu(tmp_path, defines):
    exe = _build(tmp_path, defines)
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "small.tns")], capture_output=True,
                       text=True, timeout=300)
    if r.returncode == 0:  # a GPU is present: the calls ran on it
        assert "fit " in r.stdout and "err " in r.stdout
        return
    assert "create_alto failed" in r.stderr and "no" in r.stderr.lower()


@pytest.mark.gpu
@pytest.mark.parametrize("defines", DEFINES)
def test_reference_style_program_runs_on_gpu(tmp_path, defines):
    """med.tns (2425 x 10816 x 29567, 100000 nonzeros: 41, 9 and 3 per row):
    with GENERATE_CR_BITS only mode 2 (at most 4 per row) takes a bit."""
    exe = _build(tmp_path, defines)
    print("drop-in program", defines, flush=True)
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "med.tns")], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    cr = "-DGENERATE_CR_BITS" in defines
    assert ("cr modes: 0 0 1 union 1" if cr else "cr modes: 0 0 0 union 0") in r.stdout
    assert "fit " in r.stdout and "err " in r.stdout
