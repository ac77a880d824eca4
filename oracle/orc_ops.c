/*
 * orc_ops.c -- SBP stencil operators of the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Restates src/StencilOperatorImpl.f90 of magudi:
 *   setupOperator      :1111-2193 (SBP 4-8 tables :1586-1959, right mirror :2181-2191)
 *   getAdjointOperator :2285-2372
 *   applyOperator_{1,2,3} + applyOperatorAtInteriorPoints_{1,2,3} :35-457
 *   applyOperatorNorm / NormInverse :838-1109
 * and the periodic ghost fill of src/MPIHelperImpl.f90:113-389 (single rank).
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "magudi_oracle.h"

#define B1(op, c, m) ((op)->b1[((m) - 1) * ORC_MAXB + ((c) - 1)])
#define B2(op, c, m) ((op)->b2[((m) - 1) * ORC_MAXB + ((c) - 1)])
#define INT(op, k) ((op)->interior[(k) + ORC_MAXB])

static void fill_norm48(orc_op* op) {
  /* H boundary weights of the SBP 4-8 operator (src/StencilOperatorImpl.f90:1600-1607) */
  const double h[8] = {1498139.0 / 5080320.0, 1107307.0 / 725760.0, 20761.0 / 80640.0,
                       1304999.0 / 725760.0,  299527.0 / 725760.0,  103097.0 / 80640.0,
                       670091.0 / 725760.0,   5127739.0 / 5080320.0};
  for (int i = 0; i < 8; ++i) op->norm[i] = h[i];
  op->nnorm = 8;
}

static void setup_first48(orc_op* op) {
  /* src/StencilOperatorImpl.f90:1586-1743 */
  op->sym = 1; op->iw = 9; op->bw = 12; op->bd = 8;
  INT(op, 1) = 4.0 / 5.0; INT(op, 2) = -1.0 / 5.0; INT(op, 3) = 4.0 / 105.0; INT(op, 4) = -1.0 / 280.0;
  for (int k = 1; k <= 4; ++k) INT(op, -k) = -INT(op, k);
  fill_norm48(op);
  const double x1 = 541.0 / 1000.0, x2 = -27.0 / 400.0, x3 = 187.0 / 250.0;

  B1(op, 1, 1) = -2540160.0 / 1498139.0;
  B1(op, 2, 1) = 9.0 * (2257920.0 * x1 + 11289600.0 * x2 + 22579200.0 * x3 - 15849163.0) / 5992556.0;
  B1(op, 3, 1) = 3.0 * (-33868800.0 * x1 - 162570240.0 * x2 - 304819200.0 * x3 + 235236677.0) / 5992556.0;
  B1(op, 4, 1) = (609638400.0 * x1 + 2743372800.0 * x2 + 4572288000.0 * x3 - 3577778591.0) / 17977668.0;
  B1(op, 5, 1) = 3.0 * (-16934400.0 * x1 - 67737600.0 * x2 - 84672000.0 * x3 + 67906303.0) / 1498139.0;
  B1(op, 6, 1) = 105.0 * (967680.0 * x1 + 2903040.0 * x2 - 305821.0) / 5992556.0;
  B1(op, 7, 1) = 49.0 * (-1244160.0 * x1 + 18662400.0 * x3 - 13322233.0) / 17977668.0;
  B1(op, 8, 1) = 3.0 * (-6773760.0 * x2 - 33868800.0 * x3 + 24839327.0) / 5992556.0;

  B1(op, 1, 2) = 9.0 * (-2257920.0 * x1 - 11289600.0 * x2 - 22579200.0 * x3 + 15849163.0) / 31004596.0;
  B1(op, 2, 2) = 0.0;
  B1(op, 3, 2) = 3.0 * (7257600.0 * x1 + 33868800.0 * x2 + 60963840.0 * x3 - 47167457.0) / 2214614.0;
  B1(op, 4, 2) = 3.0 * (-9676800.0 * x1 - 42336000.0 * x2 - 67737600.0 * x3 + 53224573.0) / 1107307.0;
  B1(op, 5, 2) = 7.0 * (55987200.0 * x1 + 217728000.0 * x2 + 261273600.0 * x3 - 211102099.0) / 13287684.0;
  B1(op, 6, 2) = 3.0 * (-11612160.0 * x1 - 33868800.0 * x2 + 3884117.0) / 2214614.0;
  B1(op, 7, 2) = 150.0 * (24192.0 * x1 - 338688.0 * x3 + 240463.0) / 1107307.0;
  B1(op, 8, 2) = (152409600.0 * x2 + 731566080.0 * x3 - 536324953.0) / 46506894.0;

  B1(op, 1, 3) = (33868800.0 * x1 + 162570240.0 * x2 + 304819200.0 * x3 - 235236677.0) / 1743924.0;
  B1(op, 2, 3) = (-7257600.0 * x1 - 33868800.0 * x2 - 60963840.0 * x3 + 47167457.0) / 124566.0;
  B1(op, 3, 3) = 0.0;
  B1(op, 4, 3) = (24192000.0 * x1 + 101606400.0 * x2 + 152409600.0 * x3 - 120219461.0) / 124566.0;
  B1(op, 5, 3) = (-72576000.0 * x1 - 270950400.0 * x2 - 304819200.0 * x3 + 249289259.0) / 249132.0;
  B1(op, 6, 3) = 9.0 * (806400.0 * x1 + 2257920.0 * x2 - 290167.0) / 41522.0;
  B1(op, 7, 3) = 6.0 * (-134400.0 * x1 + 1693440.0 * x3 - 1191611.0) / 20761.0;
  B1(op, 8, 3) = 5.0 * (-2257920.0 * x2 - 10160640.0 * x3 + 7439833.0) / 290654.0;

  B1(op, 1, 4) = (-609638400.0 * x1 - 2743372800.0 * x2 - 4572288000.0 * x3 + 3577778591.0) / 109619916.0;
  B1(op, 2, 4) = 3.0 * (9676800.0 * x1 + 42336000.0 * x2 + 67737600.0 * x3 - 53224573.0) / 1304999.0;
  B1(op, 3, 4) = 3.0 * (-24192000.0 * x1 - 101606400.0 * x2 - 152409600.0 * x3 + 120219461.0) / 2609998.0;
  B1(op, 4, 4) = 0.0;
  B1(op, 5, 4) = 9.0 * (16128000.0 * x1 + 56448000.0 * x2 + 56448000.0 * x3 - 47206049.0) / 5219996.0;
  B1(op, 6, 4) = 3.0 * (-19353600.0 * x1 - 50803200.0 * x2 + 7628371.0) / 2609998.0;
  B1(op, 7, 4) = 2.0 * (10886400.0 * x1 - 114307200.0 * x3 + 79048289.0) / 3914997.0;
  B1(op, 8, 4) = 75.0 * (1354752.0 * x2 + 5419008.0 * x3 - 3952831.0) / 18269986.0;

  B1(op, 1, 5) = 3.0 * (16934400.0 * x1 + 67737600.0 * x2 + 84672000.0 * x3 - 67906303.0) / 2096689.0;
  B1(op, 2, 5) = 7.0 * (-55987200.0 * x1 - 217728000.0 * x2 - 261273600.0 * x3 + 211102099.0) / 3594324.0;
  B1(op, 3, 5) = 3.0 * (72576000.0 * x1 + 270950400.0 * x2 + 304819200.0 * x3 - 249289259.0) / 1198108.0;
  B1(op, 4, 5) = 9.0 * (-16128000.0 * x1 - 56448000.0 * x2 - 56448000.0 * x3 + 47206049.0) / 1198108.0;
  B1(op, 5, 5) = 0.0;
  B1(op, 6, 5) = 105.0 * (414720.0 * x1 + 967680.0 * x2 - 165527.0) / 1198108.0;
  B1(op, 7, 5) = 15.0 * (-967680.0 * x1 + 6773760.0 * x3 - 4472029.0) / 1198108.0;
  B1(op, 8, 5) = (-304819200.0 * x2 - 914457600.0 * x3 + 657798011.0) / 25160268.0;
  B1(op, 9, 5) = -2592.0 / 299527.0;

  B1(op, 1, 6) = 5.0 * (-967680.0 * x1 - 2903040.0 * x2 + 305821.0) / 1237164.0;
  B1(op, 2, 6) = (11612160.0 * x1 + 33868800.0 * x2 - 3884117.0) / 618582.0;
  B1(op, 3, 6) = 9.0 * (-806400.0 * x1 - 2257920.0 * x2 + 290167.0) / 206194.0;
  B1(op, 4, 6) = (19353600.0 * x1 + 50803200.0 * x2 - 7628371.0) / 618582.0;
  B1(op, 5, 6) = 35.0 * (-414720.0 * x1 - 967680.0 * x2 + 165527.0) / 1237164.0;
  B1(op, 6, 6) = 0.0;
  B1(op, 7, 6) = 80640.0 * x1 / 103097.0;
  B1(op, 8, 6) = 80640.0 * x2 / 103097.0;
  B1(op, 9, 6) = 3072.0 / 103097.0;
  B1(op, 10, 6) = -288.0 / 103097.0;

  B1(op, 1, 7) = 7.0 * (1244160.0 * x1 - 18662400.0 * x3 + 13322233.0) / 8041092.0;
  B1(op, 2, 7) = 150.0 * (-24192.0 * x1 + 338688.0 * x3 - 240463.0) / 670091.0;
  B1(op, 3, 7) = 54.0 * (134400.0 * x1 - 1693440.0 * x3 + 1191611.0) / 670091.0;
  B1(op, 4, 7) = 2.0 * (-10886400.0 * x1 + 114307200.0 * x3 - 79048289.0) / 2010273.0;
  B1(op, 5, 7) = 15.0 * (967680.0 * x1 - 6773760.0 * x3 + 4472029.0) / 2680364.0;
  B1(op, 6, 7) = -725760.0 * x1 / 670091.0;
  B1(op, 7, 7) = 0.0;
  B1(op, 8, 7) = 725760.0 * x3 / 670091.0;
  B1(op, 9, 7) = -145152.0 / 670091.0;
  B1(op, 10, 7) = 27648.0 / 670091.0;
  B1(op, 11, 7) = -2592.0 / 670091.0;

  B1(op, 1, 8) = 3.0 * (6773760.0 * x2 + 33868800.0 * x3 - 24839327.0) / 20510956.0;
  B1(op, 2, 8) = (-152409600.0 * x2 - 731566080.0 * x3 + 536324953.0) / 30766434.0;
  B1(op, 3, 8) = 45.0 * (2257920.0 * x2 + 10160640.0 * x3 - 7439833.0) / 10255478.0;
  B1(op, 4, 8) = 75.0 * (-1354752.0 * x2 - 5419008.0 * x3 + 3952831.0) / 10255478.0;
  B1(op, 5, 8) = (304819200.0 * x2 + 914457600.0 * x3 - 657798011.0) / 61532868.0;
  B1(op, 6, 8) = -5080320.0 * x2 / 5127739.0;
  B1(op, 7, 8) = -5080320.0 * x3 / 5127739.0;
  B1(op, 8, 8) = 0.0;
  B1(op, 9, 8) = 4064256.0 / 5127739.0;
  B1(op, 10, 8) = -1016064.0 / 5127739.0;
  B1(op, 11, 8) = 193536.0 / 5127739.0;
  B1(op, 12, 8) = -18144.0 / 5127739.0;
}

static void set_row(orc_op* op, int m, int c0, int cnt, const double* v, double scale) {
  for (int i = 0; i < cnt; ++i) B1(op, c0 + i, m) = v[i] / scale;
}

static void setup_second48(orc_op* op) {
  /* src/StencilOperatorImpl.f90:1745-1840 */
  op->sym = 0; op->iw = 9; op->bw = 12; op->bd = 8;
  INT(op, 0) = -205.0 / 72.0; INT(op, 1) = 8.0 / 5.0; INT(op, 2) = -1.0 / 5.0;
  INT(op, 3) = 8.0 / 315.0; INT(op, 4) = -1.0 / 560.0;
  for (int k = 1; k <= 4; ++k) INT(op, -k) = INT(op, k);
  fill_norm48(op);
  const double r1[7] = {4870382994799.0 / 1358976868290.0, -893640087518.0 / 75498714905.0,
                        926594825119.0 / 60398971924.0,    -1315109406200.0 / 135897686829.0,
                        39126983272.0 / 15099742981.0,     12344491342.0 / 75498714905.0,
                        -451560522577.0 / 2717953736580.0};
  const double r2[8] = {333806012194.0 / 390619153855.0, -154646272029.0 / 111605472530.0,
                        1168338040.0 / 33481641759.0,    82699112501.0 / 133926567036.0,
                        -171562838.0 / 11160547253.0,    -28244698346.0 / 167408208795.0,
                        11904122576.0 / 167408208795.0,  -2598164715.0 / 312495323084.0};
  const double r3[8] = {7838984095.0 / 52731029988.0,   1168338040.0 / 5649753213.0,
                        -88747895.0 / 144865467.0,      423587231.0 / 627750357.0,
                        -43205598281.0 / 22599012852.0, 4876378562.0 / 1883251071.0,
                        -5124426509.0 / 3766502142.0,   10496900965.0 / 39548272491.0};
  const double r4[8] = {-94978241528.0 / 828644350023.0,  82699112501.0 / 157837019052.0,
                        1270761693.0 / 13153084921.0,     -167389605005.0 / 118377764289.0,
                        48242560214.0 / 39459254763.0,    -31673996013.0 / 52612339684.0,
                        43556319241.0 / 118377764289.0,   -44430275135.0 / 552429566682.0};
  const double r5[9] = {1455067816.0 / 21132528431.0,   -171562838.0 / 3018932633.0,
                        -43205598281.0 / 36227191596.0, 48242560214.0 / 9056797899.0,
                        -52276055645.0 / 6037865266.0,  57521587238.0 / 9056797899.0,
                        -80321706377.0 / 36227191596.0, 8078087158.0 / 21132528431.0,
                        -1296.0 / 299527.0};
  const double r6[10] = {10881504334.0 / 327321118845.0,  -28244698346.0 / 140280479505.0,
                         4876378562.0 / 9352031967.0,     -10557998671.0 / 12469375956.0,
                         57521587238.0 / 28056095901.0,   -278531401019.0 / 93520319670.0,
                         73790130002.0 / 46760159835.0,   -137529995233.0 / 785570685228.0,
                         2048.0 / 103097.0,               -144.0 / 103097.0};
  const double r7[11] = {-135555328849.0 / 8509847458140.0, 11904122576.0 / 101307707835.0,
                         -5124426509.0 / 13507694378.0,     43556319241.0 / 60784624701.0,
                         -80321706377.0 / 81046166268.0,    73790130002.0 / 33769235945.0,
                         -950494905688.0 / 303923123505.0,  239073018673.0 / 141830790969.0,
                         -145152.0 / 670091.0,              18432.0 / 670091.0,
                         -1296.0 / 670091.0};
  const double r8[11] = {-2598164715.0 / 206729925524.0, 10496900965.0 / 155047444143.0,
                         -44430275135.0 / 310094888286.0, 425162482.0 / 2720130599.0,
                         -137529995233.0 / 620189776572.0, 239073018673.0 / 155047444143.0,
                         -144648000000.0 / 51682481381.0,  8128512.0 / 5127739.0,
                         -1016064.0 / 5127739.0,           129024.0 / 5127739.0,
                         -9072.0 / 5127739.0};
  set_row(op, 1, 1, 7, r1, 1.0); set_row(op, 2, 1, 8, r2, 1.0);
  set_row(op, 3, 1, 8, r3, 1.0); set_row(op, 4, 1, 8, r4, 1.0);
  set_row(op, 5, 1, 9, r5, 1.0); set_row(op, 6, 1, 10, r6, 1.0);
  set_row(op, 7, 1, 11, r7, 1.0); set_row(op, 8, 2, 11, r8, 1.0);
}

static void setup_composite48(orc_op* op) {
  /* src/StencilOperatorImpl.f90:1842-1926: 8th difference / 256 */
  op->sym = 0; op->iw = 9; op->bw = 12; op->bd = 8;
  const double c[5] = {-70.0, 56.0, -28.0, 8.0, -1.0};
  for (int k = 0; k <= 4; ++k) INT(op, k) = c[k];
  for (int k = 1; k <= 4; ++k) INT(op, -k) = INT(op, k);
  for (int k = -4; k <= 4; ++k) INT(op, k) = INT(op, k) / 256.0;
  fill_norm48(op);
  const double r1[5] = {-15240960.0 / 1498139.0, 60963840.0 / 1498139.0, -91445760.0 / 1498139.0,
                        60963840.0 / 1498139.0, -15240960.0 / 1498139.0};
  const double r2[6] = {8709120.0 / 1107307.0, -35562240.0 / 1107307.0, 55157760.0 / 1107307.0,
                        -39191040.0 / 1107307.0, 11612160.0 / 1107307.0, -725760.0 / 1107307.0};
  const double r3[7] = {-1451520.0 / 20761.0, 6128640.0 / 20761.0, -10080000.0 / 20761.0,
                        8064000.0 / 20761.0,  -3225600.0 / 20761.0, 645120.0 / 20761.0,
                        -80640.0 / 20761.0};
  const double r4[8] = {8709120.0 / 1304999.0,  -39191040.0 / 1304999.0, 72576000.0 / 1304999.0,
                        -73301760.0 / 1304999.0, 46448640.0 / 1304999.0, -20321280.0 / 1304999.0,
                        5806080.0 / 1304999.0,  -725760.0 / 1304999.0};
  const double r5[9] = {-2177280.0 / 299527.0,  11612160.0 / 299527.0, -29030400.0 / 299527.0,
                        46448640.0 / 299527.0,  -52254720.0 / 299527.0, 40642560.0 / 299527.0,
                        -20321280.0 / 299527.0, 5806080.0 / 299527.0,  -725760.0 / 299527.0};
  const double r6[9] = {-80640.0 / 103097.0,   645120.0 / 103097.0,  -2257920.0 / 103097.0,
                        4515840.0 / 103097.0,  -5644800.0 / 103097.0, 4515840.0 / 103097.0,
                        -2257920.0 / 103097.0, 645120.0 / 103097.0,  -80640.0 / 103097.0};
  const double r7[9] = {-725760.0 / 670091.0,   5806080.0 / 670091.0,  -20321280.0 / 670091.0,
                        40642560.0 / 670091.0,  -50803200.0 / 670091.0, 40642560.0 / 670091.0,
                        -20321280.0 / 670091.0, 5806080.0 / 670091.0,  -725760.0 / 670091.0};
  const double r8[9] = {-5080320.0 / 5127739.0,    40642560.0 / 5127739.0,  -142248960.0 / 5127739.0,
                        284497920.0 / 5127739.0,   -355622400.0 / 5127739.0, 284497920.0 / 5127739.0,
                        -142248960.0 / 5127739.0,  40642560.0 / 5127739.0,  -5080320.0 / 5127739.0};
  set_row(op, 1, 1, 5, r1, 1.0); set_row(op, 2, 1, 6, r2, 1.0);
  set_row(op, 3, 1, 7, r3, 1.0); set_row(op, 4, 1, 8, r4, 1.0);
  set_row(op, 5, 1, 9, r5, 1.0); set_row(op, 6, 2, 9, r6, 1.0);
  set_row(op, 7, 3, 9, r7, 1.0); set_row(op, 8, 4, 9, r8, 1.0);
  for (int m = 1; m <= 8; ++m)
    for (int cc = 1; cc <= 12; ++cc) B1(op, cc, m) = B1(op, cc, m) / 256.0;
}

static void setup_diss48(orc_op* op) {
  /* src/StencilOperatorImpl.f90:1928-1940 */
  op->sym = 0; op->iw = 5; op->bw = 5; op->bd = 2;
  INT(op, 0) = 6.0; INT(op, 1) = -4.0; INT(op, 2) = 1.0;
  INT(op, -1) = INT(op, 1); INT(op, -2) = INT(op, 2);
  for (int m = 1; m <= 2; ++m)
    for (int k = -2; k <= 2; ++k) B1(op, k + 3, m) = INT(op, k);
  op->nnorm = 2; op->norm[0] = op->norm[1] = 1.0;
}

static void setup_dissT48(orc_op* op) {
  /* src/StencilOperatorImpl.f90:1942-1959 */
  op->sym = 0; op->iw = 5; op->bw = 7; op->bd = 5;
  INT(op, 0) = 6.0; INT(op, 1) = -4.0; INT(op, 2) = 1.0;
  INT(op, -1) = INT(op, 1); INT(op, -2) = INT(op, 2);
  for (int cc = 1; cc <= 3; ++cc)
    for (int m = 1; m <= 5; ++m) B1(op, cc, m) = INT(op, m - 3);
  for (int m = 2; m <= 5; ++m) B1(op, 4, m) = INT(op, m - 4);
  for (int m = 3; m <= 5; ++m) B1(op, 5, m) = INT(op, m - 5);
  for (int m = 4; m <= 5; ++m) B1(op, 6, m) = INT(op, m - 6);
  B1(op, 7, 5) = INT(op, -2);
  op->nnorm = 5;
  for (int i = 0; i < 5; ++i) op->norm[i] = 1.0;
}

static void fill_right(orc_op* op) {
  /* right-boundary mirror, src/StencilOperatorImpl.f90:2181-2191 */
  const double s = op->sym == 1 ? -1.0 : 1.0;
  for (int m = 1; m <= op->bd; ++m)
    for (int cc = 1; cc <= op->bw; ++cc) B2(op, cc, m) = s * B1(op, op->bw + 1 - cc, m);
}

int orc_op_setup(orc_op* op, int scheme, int kind) {
  memset(op, 0, sizeof *op);
  if (scheme != ORC_SBP48) return -1;
  switch (kind) {
    case ORC_FIRST: setup_first48(op); break;
    case ORC_SECOND: setup_second48(op); break;
    case ORC_DISS: setup_diss48(op); break;
    case ORC_DISS_T: setup_dissT48(op); break;
    case ORC_COMPOSITE: setup_composite48(op); break;
    default: return -1;
  }
  fill_right(op);
  return 0;
}

void orc_op_adjoint(const orc_op* op, orc_op* adj) {
  /* getAdjointOperator, src/StencilOperatorImpl.f90:2285-2372 */
  memset(adj, 0, sizeof *adj);
  const int hw = op->iw / 2;
  adj->sym = op->sym;
  adj->iw = op->iw;
  adj->bd = op->bw;
  adj->bw = op->bw + hw;
  for (int i = -hw; i <= hw; ++i) INT(adj, i) = INT(op, -i);
  adj->nnorm = op->bd;
  for (int i = 0; i < op->bd; ++i) adj->norm[i] = op->norm[i];
  /* transpose of the boundary block into columns 1..bd */
  for (int i = 1; i <= op->bd; ++i)
    for (int m = 1; m <= op->bw; ++m) B1(adj, i, m) = B1(op, m, i);
  /* interior rows bd+1 .. bw+hw of the original, transposed */
  for (int i = op->bd + 1; i <= op->bw + hw; ++i)
    for (int j = -hw; j <= hw; ++j) {
      if (i + j > op->bw) break;
      B1(adj, i, i + j) = INT(op, j);
    }
  const int nn = adj->nnorm;
  for (int i = 1; i <= adj->bw; ++i)       /* pre-multiply by H^-1 */
    for (int r = 1; r <= nn; ++r) B1(adj, i, r) = B1(adj, i, r) / adj->norm[r - 1];
  for (int i = 1; i <= adj->bd; ++i)       /* post-multiply by H */
    for (int c = 1; c <= nn; ++c) B1(adj, c, i) = B1(adj, c, i) * adj->norm[c - 1];
  fill_right(adj);
}

static inline int wrap_index(int i, int len, int off) {
  /* periodic ghost map of fillGhostPoints (src/MPIHelperImpl.f90:200-296), one rank */
  if (i < 0) return i + len - off;
  if (i >= len) return i - len + off;
  return i;
}

/* Apply along one pencil: in -> out (both contiguous of length len). */
static void apply_pencil(const orc_op* op, int len, int periodic, int off, const double* in,
                         double* out) {
  const int hw = op->iw / 2;
  if (len == 1) { out[0] = 0.0; return; }   /* "null matrix" for a collapsed direction */
  int is = 0, ie = len - 1;
  if (!periodic) { is = op->bd; ie = len - 1 - op->bd; }
  for (int i = is; i <= ie; ++i) {
    double s = 0.0;
    if (op->sym == 1) {
      for (int k = 1; k <= hw; ++k)
        s += INT(op, k) * (in[wrap_index(i + k, len, off)] - in[wrap_index(i - k, len, off)]);
    } else {
      for (int k = 1; k <= hw; ++k)
        s += INT(op, k) * (in[wrap_index(i + k, len, off)] + in[wrap_index(i - k, len, off)]);
      s = s + INT(op, 0) * in[i];
    }
    out[i] = s;
  }
  if (!periodic) {
    const int n = op->bw;
    for (int m = 1; m <= op->bd; ++m) {
      double s = 0.0;
      for (int c = 1; c <= n; ++c) s += B1(op, c, m) * in[c - 1];
      out[m - 1] = s;
    }
    for (int m = op->bd; m >= 1; --m) {
      double s = 0.0;
      for (int c = 1; c <= n; ++c) s += B2(op, c, m) * in[len - n + c - 1];
      out[len - m] = s;
    }
  }
}

void orc_op_apply(const orc_op* op, int dir, const int n[3], int periodicity, double* x,
                  int ncomp) {
  const long N = (long)n[0] * n[1] * n[2];
  const int len = n[dir];
  const long stride = dir == 0 ? 1 : (dir == 1 ? n[0] : (long)n[0] * n[1]);
  const int periodic = periodicity != ORC_NONE;
  const int off = periodicity == ORC_OVERLAP ? 1 : 0;
  const int na = dir == 0 ? n[1] : n[0];
  const int nb = dir == 2 ? n[1] : n[2];
  const long npencil = (long)ncomp * nb * na;
  /* pencils are independent: split over the OpenMP threads, one buffer pair each */
#pragma omp parallel
  {
    double* in = malloc(sizeof(double) * len);
    double* out = malloc(sizeof(double) * len);
#pragma omp for schedule(static)
    for (long w = 0; w < npencil; ++w) {
      const int l = (int)(w / ((long)nb * na));
      const int b = (int)((w / na) % nb);
      const int a = (int)(w % na);
      long base;
      if (dir == 0) base = (long)n[0] * (a + (long)n[1] * b);
      else if (dir == 1) base = a + (long)n[0] * n[1] * b;
      else base = a + (long)n[0] * b;
      double* xl = x + l * N + base;
      for (int i = 0; i < len; ++i) in[i] = xl[i * stride];
      apply_pencil(op, len, periodic, off, in, out);
      for (int i = 0; i < len; ++i) xl[i * stride] = out[i];
    }
    free(in);
    free(out);
  }
}

void orc_op_apply_norm(const orc_op* op, int dir, const int n[3], int periodic, double* x,
                       int ncomp, int inverse) {
  /* applyOperatorNorm/NormInverse (src/StencilOperatorImpl.f90:838-1109) */
  if (periodic) return;
  const long N = (long)n[0] * n[1] * n[2];
  for (int l = 0; l < ncomp; ++l)
#pragma omp parallel for schedule(static)
    for (long p = 0; p < N; ++p) {
      int idx = dir == 0 ? (int)(p % n[0]) : (dir == 1 ? (int)((p / n[0]) % n[1])
                                                        : (int)(p / ((long)n[0] * n[1])));
      int len = n[dir];
      int w = -1;
      if (idx < op->bd) w = idx;
      else if (idx >= len - op->bd) w = len - 1 - idx;
      if (w < 0) continue;
      if (inverse) x[l * N + p] = x[l * N + p] / op->norm[w];
      else x[l * N + p] = op->norm[w] * x[l * N + p];
    }
}

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
