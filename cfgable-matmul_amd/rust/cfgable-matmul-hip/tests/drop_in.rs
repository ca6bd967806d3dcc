// The reference's own call forms, unchanged, on small shapes against naive loops.  Needs an
// MI355X and libcfgmm.so (CFGMM_LIB_DIR or the in-tree build).
use cfgable_matmul::impl_aodmao2rho::aodmao2rho_anyway;
use cfgable_matmul::impl_matmul::matmul_anyway_full;
use cfgable_matmul::structs::MatmulLoops;

fn naive_gemm(c: &mut [f64], a: &[f64], b: &[f64], m: usize, n: usize, k: usize, lda: usize, ldb: usize, ldc: usize,
              transa: bool, transb: bool) {
    for i in 0..m {
        for j in 0..n {
            let mut s = 0.0;
            for p in 0..k {
                let x = if transa { a[p * lda + i] } else { a[i * lda + p] };
                let y = if transb { b[j * ldb + p] } else { b[p * ldb + j] };
                s += x * y;
            }
            c[i * ldc + j] += s;
        }
    }
}

fn rel(x: &[f64], y: &[f64]) -> f64 {
    let num: f64 = x.iter().zip(y).map(|(a, b)| (a - b) * (a - b)).sum();
    let den: f64 = y.iter().map(|b| b * b).sum();
    (num / den).sqrt()
}

#[test]
fn gemm_both_turbofish_forms() {
    let (m, n, k) = (131, 97, 203);
    for &(ta, tb) in &[(false, false), (false, true), (true, false), (true, true)] {
        let lda = if ta { m } else { k };
        let ldb = if tb { k } else { n };
        let a: Vec<f64> = (0..m * k).map(|x| (x as f64).sin()).collect();
        let b: Vec<f64> = (0..k * n).map(|x| (x as f64).cos()).collect();
        let mut want = vec![0.0; m * n];
        naive_gemm(&mut want, &a, &b, m, n, k, lda, ldb, n, ta, tb);
        let mut c = vec![0.0; m * n];
        MatmulLoops::<f64, 252, 512, 240, 14, 2, 8>::matmul_loop_parallel_mnk_pack_a(&mut c, &a, &b, m, n, k, lda,
                                                                                     ldb, n, ta, tb);
        assert!(rel(&c, &want) <= 1e-12);
        let mut c2 = vec![0.0; m * n];
        MatmulLoops::<f64, 252, 512, 240, 14, 2, 8, 2360>::matmul_loop_macro_mb(&mut c2, &a, &b, m, n, k, lda, ldb,
                                                                                n, ta, tb);
        assert!(rel(&c2, &want) <= 1e-12);
        let mut c3 = vec![0.0; m * n];
        matmul_anyway_full(&mut c3, &a, &b, m, n, k, lda, ldb, n, ta, tb);
        assert!(rel(&c3, &want) <= 1e-12);
    }
}

#[test]
fn rho_reference_form() {
    let (nao, ngrid, nset) = (90usize, 250usize, 2usize);
    const BLKSIZE: usize = 48;
    let nblk = ngrid.div_ceil(BLKSIZE);
    let pattern = [true, false, true, false, false, false, false, false, false, false, false, true, true, false,
                   false, false, false, false, true, false];
    let tab: Vec<bool> = (0..nao * nblk).map(|i| pattern[i % 20]).collect();
    let mut lhs: Vec<f64> = (0..nao * ngrid).map(|x| (x as f64).sin()).collect();
    let mut rhs: Vec<f64> = (0..nao * ngrid).map(|x| (x as f64).cos()).collect();
    let dm: Vec<f64> = (0..nset * nao * nao).map(|x| (x as f64 * 0.38 + 1.2).sin()).collect();
    for p in 0..nao {
        for blk in 0..nblk {
            if !tab[p * nblk + blk] {
                for g in blk * BLKSIZE..((blk + 1) * BLKSIZE).min(ngrid) {
                    lhs[p * ngrid + g] = 0.0;
                    rhs[p * ngrid + g] = 0.0;
                }
            }
        }
    }
    let mut want = vec![0.0; nset * ngrid];
    for s in 0..nset {
        for g in 0..ngrid {
            for u in 0..nao {
                for v in 0..nao {
                    want[s * ngrid + g] += lhs[u * ngrid + g] * dm[s * nao * nao + u * nao + v] * rhs[v * ngrid + g];
                }
            }
        }
    }
    let mut rho = vec![0.0; nset * ngrid];
    aodmao2rho_anyway(&mut rho, &dm, &lhs, &rhs, [nao, ngrid], [nao, ngrid], nset, &tab, nblk);
    assert!(rel(&rho, &want) <= 1e-12);
    let mut rho2 = vec![0.0; nset * ngrid];
    MatmulLoops::<f64, 252, 512, 240, 14, 2, 8>::aodmao2rho_loop_parallel::<BLKSIZE>(
        &mut rho2, &dm, &lhs, &rhs, [nao, ngrid], [nao, ngrid], nset, &tab, nblk);
    assert!(rel(&rho2, &want) <= 1e-12);
}

#[test]
#[should_panic(expected = "transb=true is not implemented")]
fn masked_transb_panics_like_the_reference() {
    let (m, n, k) = (4, 5, 6);
    let tab = vec![true; k];
    let mut c = vec![0.0; m * n];
    cfgable_matmul::impl_matmul_non0tab::matmul_anyway_full_non0tab(
        &mut c, &vec![0.0; m * k], &vec![0.0; n * k], m, n, k, k, k, n, false, true, &tab, 1);
}
