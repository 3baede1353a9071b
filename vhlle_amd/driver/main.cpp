// vhlle: command-line driver over the GPU hydro step, with the reference's
// params-file interface (src/main.cpp:70-156, 226-243, 399-428).
//
// Same keys, defaults and CLI (-params, -system, -ISinput, -outputDir) as the
// reference; keys the reference does not know are echoed "UUU", comment lines
// ('!') "CCC".  eosType 0 reads the tabulated EoS eosFile (default
// eos/Laine_nf3.dat, the reference's path; columns e, p, T) as EoSs does.
// eosType 1 reads eos/chiraleos.dat and eos/chiralsmall.dat (EoSChiral);
// eosTypeHadron 0 reads eos/eosHadronLog.dat (EoSHadron) as the EoS of the
// freeze-out outputs, as the reference does.
// Added keys: eosType 5 = ideal gas p=e/3; eosTypeHadron -1 = the hydro EoS
// also for the freeze-out outputs; eosFile; icModel 10 = smooth
// Gaussian state (icE0, icSigmaX, icSigmaY, icEtaPlateau, icEtaSigma);
// gubserScale (energy scale of icModel 4); zetaSPeakT, zetaSWidthMinus,
// zetaSWidthPlus (zetaSparam 5); dumpState 1 writes the final conserved
// state to outputDir/state.bin (7 x N doubles, C-ABI layout); precision
// (0 FP64, 1 FP32 fluxes), c2pMode (0 reference, 1 learned, 2 closed form,
// 3 Newton from v = 0.5), c2pModel (a file written by -trainC2P);
// stopOnEmptySurface (1, the reference's loop condition: stop at the first
// step without freeze-out elements, which for nz < 5 is the first step).
// vorticity 1 writes beta.dat next to freezeout.dat;
// VTK_output_values (e.g. eps,T,v,pi) writes VTK files before every step.
// tauGridResize: once tau reaches it the grid doubles its transverse extent
// (expandGrid2x, src/main.cpp:246-271).
// `vhlle [-params file] -trainC2P file [-seed S]` harvests 4 M conserved ->
// primitive pairs from Gubser runs with the params' EoS (ideal gas without
// -params), trains the con2prim MLP on the GPU and writes it.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <fstream>
#include <functional>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../host/vhlle.h"

using namespace std;

namespace {

int nx{100}, ny{100}, nz{100}, eosType{1}, etaSparam{0}, zetaSparam{0}, eosTypeHadron{0};
string eosFile{"eos/Laine_nf3.dat"};
bool freezeoutOnly{false}, freezeoutExtend{false}, vorticityOn{false};
double xmin{-5.0}, xmax{5.0}, ymin{-5.0}, ymax{5.0}, etamin{-5.0}, etamax{5.0}, tau0{1.0},
    tauMax{20.0}, tauResize{4.0}, dtau{0.05}, etaS{0.08}, zetaS{0.0}, eCrit{0.5},
    etaSEpsilonMin{5.}, al{0.}, ah{0.}, aRho{0.}, T0{0.15}, etaSMin{0.08}, etaSShiftMuB{0.},
    etaSScaleMuB{0.}, zetaSPeakEpsilon{5.}, zetaSScaleBeta{0.103}, zetaSSigmaMinus{0.1},
    zetaSSigmaPlus{0.1}, epsilon0, Rgt{1.0}, Rgz{1.0}, impactPar, s0ScaleFactor;
string collSystem, outputDir{"data"}, isInputFile, vtk_values{""};
int icModel{1}, glauberVariable{1}, smoothingType{0};
// added keys
double gubserScale{1.0}, icE0{50.0}, icSigmaX{2.5}, icSigmaY{3.0}, icEtaPlateau{1.5},
    icEtaSigma{1.0}, zetaSPeakT{0.170}, zetaSWidthMinus{0.010}, zetaSWidthPlus{0.030};
int dumpState{0}, precision{0}, c2pMode{0}, stopOnEmptySurface{1};
string c2pModel;

void readParameters(const char *parFile) {
  ifstream fin(parFile);
  if (!fin.is_open()) {
    cout << "cannot open parameters file " << parFile << endl;
    exit(1);
  }
  cout << "vhlle: reading parameters from " << parFile << endl;
  auto I = [](int &v) { return [&v](const string &s) { v = atoi(s.c_str()); }; };
  auto D = [](double &v) { return [&v](const string &s) { v = atof(s.c_str()); }; };
  auto B = [](bool &v) { return [&v](const string &s) { v = atoi(s.c_str()); }; };
  map<string, function<void(const string &)>> h = {
      {"eosType", I(eosType)}, {"eosTypeHadron", I(eosTypeHadron)}, {"nx", I(nx)},
      {"ny", I(ny)}, {"nz", I(nz)}, {"icModel", I(icModel)}, {"glauberVar", I(glauberVariable)},
      {"xmin", D(xmin)}, {"xmax", D(xmax)}, {"ymin", D(ymin)}, {"ymax", D(ymax)},
      {"etamin", D(etamin)}, {"etamax", D(etamax)}, {"tau0", D(tau0)}, {"tauMax", D(tauMax)},
      {"tauGridResize", D(tauResize)}, {"dtau", D(dtau)}, {"e_crit", D(eCrit)},
      {"etaS", D(etaS)}, {"zetaS", D(zetaS)}, {"zetaSparam", I(zetaSparam)},
      {"zetaSScaleBeta", D(zetaSScaleBeta)}, {"zetaSPeakEpsilon", D(zetaSPeakEpsilon)},
      {"zetaSSigmaMinus", D(zetaSSigmaMinus)}, {"zetaSSigmaPlus", D(zetaSSigmaPlus)},
      {"epsilon0", D(epsilon0)}, {"Rg", D(Rgt)}, {"Rgz", D(Rgz)}, {"impactPar", D(impactPar)},
      {"s0ScaleFactor", D(s0ScaleFactor)},
      {"VTK_output_values", [](const string &s) { vtk_values = s; }},
      {"etaSparam", I(etaSparam)}, {"aRho", D(aRho)}, {"ah", D(ah)}, {"al", D(al)},
      {"T0", D(T0)}, {"etaSEpsilonMin", D(etaSEpsilonMin)}, {"etaSMin", D(etaSMin)},
      {"etaSShiftMuB", D(etaSShiftMuB)}, {"etaSScaleMuB", D(etaSScaleMuB)},
      {"freezeoutOnly", B(freezeoutOnly)}, {"freezeoutExtend", B(freezeoutExtend)},
      {"vorticity", B(vorticityOn)}, {"smoothingType", I(smoothingType)},
      // added by this build
      {"gubserScale", D(gubserScale)}, {"icE0", D(icE0)}, {"icSigmaX", D(icSigmaX)},
      {"icSigmaY", D(icSigmaY)}, {"icEtaPlateau", D(icEtaPlateau)},
      {"icEtaSigma", D(icEtaSigma)}, {"zetaSPeakT", D(zetaSPeakT)},
      {"zetaSWidthMinus", D(zetaSWidthMinus)}, {"zetaSWidthPlus", D(zetaSWidthPlus)},
      {"dumpState", I(dumpState)}, {"precision", I(precision)}, {"c2pMode", I(c2pMode)},
      {"c2pModel", [](const string &s) { c2pModel = s; }},
      {"eosFile", [](const string &s) { eosFile = s; }},
      {"stopOnEmptySurface", I(stopOnEmptySurface)},
  };
  while (fin.good()) {
    string line;
    getline(fin, line);
    istringstream sl(line);
    string name, value;
    sl >> name >> value;
    if (name.empty()) continue;
    auto it = h.find(name);
    if (it != h.end())
      it->second(value);
    else if (name[0] == '!')
      cout << "CCC " << sl.str() << endl;
    else
      cout << "UUU " << sl.str() << endl;
  }
  auto borders = [](double mn, double mx, const char *x) {
    if (mn >= mx) {
      cerr << "FATAL: The hydrodynamic grid is not set up correctly! \n"
           << x << "min >= " << x << "max \n";
      exit(1);
    }
  };
  borders(xmin, xmax, "x");
  borders(ymin, ymax, "y");
  borders(etamin, etamax, "eta");
}

void readCommandLine(int argc, char **argv) {
  if (argc == 1) {
    cout << "no CL params - exiting.\n";
    exit(1);
  }
  for (int i = 1; i < argc - 1; i++) {
    if (strcmp(argv[i], "-system") == 0) collSystem = argv[i + 1];
    if (strcmp(argv[i], "-params") == 0) readParameters(argv[i + 1]);
    if (strcmp(argv[i], "-ISinput") == 0) isInputFile = argv[i + 1];
    if (strcmp(argv[i], "-outputDir") == 0) outputDir = argv[i + 1];
  }
  cout << "vhlle: command line parameters are:\n"
       << "collision system:  " << collSystem << "\nini.state input:  " << isInputFile
       << "\noutput directory:  " << outputDir << endl;
}

// smooth state e = icE0 exp(-x^2/2sx^2 - y^2/2sy^2) H(eta), v = 0
void setGaussianIC(Fluid *f, EoS *eos, double tau) {
  for (int iz = 0; iz < f->getNZ(); iz++)
    for (int iy = 0; iy < f->getNY(); iy++)
      for (int ix = 0; ix < f->getNX(); ix++) {
        const double x = f->getX(ix), y = f->getY(iy);
        const double eta = f->getNZ() > 1 ? f->getZ(iz) : 0.0;
        const double a = fabs(eta);
        const double H = a < icEtaPlateau
                             ? 1.0
                             : exp(-(a - icEtaPlateau) * (a - icEtaPlateau) /
                                   (2 * icEtaSigma * icEtaSigma));
        const double e = icE0 *
                         exp(-x * x / (2 * icSigmaX * icSigmaX) - y * y / (2 * icSigmaY * icSigmaY)) *
                         H;
        Cell *c = f->getCell(ix, iy, iz);
        c->setPrimVar(eos, tau, e, 0., 0., 0., 0., 0., 0.);
        c->saveQprev();
        if (e > 0.) c->setAllM(1.);
      }
}

double totalEnergy(Fluid *f, double tau) {
  double E = 0.;
  double Q[7];
  for (int iz = 0; iz < f->getNZ(); iz++)
    for (int iy = 0; iy < f->getNY(); iy++)
      for (int ix = 0; ix < f->getNX(); ix++) {
        f->getCell(ix, iy, iz)->getQ(Q);
        E += Q[T_];
      }
  return E;
}

}  // namespace

// The grid with twice the transverse extent and the same cell counts: every
// second cell around the centre is imported (getCell clamps at the edges),
// dtau doubles, and Hydro continues on the new Fluid (src/main.cpp:246-271).
Fluid *expandGrid2x(Hydro *h, EoS *eos, EoS *eosH, TransportCoeff *trcoeff) {
  Fluid *f = h->getFluid();
  const int nx = f->getNX(), ny = f->getNY(), nz = f->getNZ();
  if (f->getX(0) + f->getX(nx - 1) > 0.001 || f->getY(0) + f->getY(ny - 1) > 0.001) {
    cout << "this grid expansion works only with symmetric min/max ranges\n";
    return f;
  }
  Fluid *fnew = new Fluid(eos, eosH, trcoeff, nx, ny, nz, 2.0 * f->getX(0), 2.0 * f->getX(nx - 1),
                          2.0 * f->getY(0), 2.0 * f->getY(ny - 1), f->getZ(0), f->getZ(nz - 1),
                          2.0 * h->getDtau(), f->geteCrit());
  if (vorticityOn) fnew->enableVorticity();  // matrices start at zero (src/main.cpp:257-265)
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++)
        fnew->getCell(ix, iy, iz)
            ->importVars(f->getCell(2 * (ix - nx / 2) + nx / 2, 2 * (iy - ny / 2) + ny / 2, iz));
  fnew->adoptOutput(f);
  h->setFluid(fnew);
  h->setDtau(2.0 * h->getDtau());
  delete f;
  return fnew;
}

// harvest + GPU training of the con2prim MLP (vh_mlp_harvest, vh_mlp_train)
int trainC2P(const char *out, unsigned long long seed, const vh_eos_t *eos) {
  const long n = 4L << 20;
  std::vector<double> Q(7 * n), prim(8 * n);
  vh_harvest_cfg_t hc{seed, 200, 2, {10.0, 1.0, 0., 0.}, 0.5, 0.99, eos};
  if (vh_mlp_harvest(&hc, n, Q.data(), prim.data()) != VH_OK) {
    cout << "harvest failed\n";
    return 1;
  }
  std::vector<float> p(VH_MLP_NPARAM);
  vh_mlp_init(seed + 1, p.data());
  vh_mlp_train_cfg_t tc{512, 8, 2e-3, 1e-6, seed + 2};
  std::vector<double> hist(tc.epochs);
  if (vh_mlp_train(&tc, Q.data(), prim.data(), n, p.data(), hist.data()) != VH_OK) {
    cout << "training failed\n";
    return 1;
  }
  for (int e = 0; e < tc.epochs; e++) cout << "epoch " << e << " loss " << hist[e] << endl;
  if (!vhlle_learn::saveC2PModel(out, p.data())) {
    cout << "cannot write " << out << endl;
    return 1;
  }
  cout << "wrote " << out << endl;
  return 0;
}

int main(int argc, char **argv) {
  const char *train = nullptr;
  unsigned long long seed = 1;
  bool params = false;
  for (int i = 1; i < argc - 1; i++) {
    if (strcmp(argv[i], "-trainC2P") == 0) train = argv[i + 1];
    if (strcmp(argv[i], "-seed") == 0) seed = strtoull(argv[i + 1], nullptr, 10);
    if (strcmp(argv[i], "-params") == 0) params = true;
  }
  if (train && !params) return trainC2P(train, seed, nullptr);
  readCommandLine(argc, argv);
  EoS *eos = nullptr;
  if (eosType == 0) {
    eos = new EoSs(eosFile, 3);  // src/main.cpp:289-290
  } else if (eosType == 1) {
    eos = new EoSChiral();  // src/main.cpp:291-292
  } else if (eosType == 5) {
    eos = new EoSIdeal();
  } else {
    cout << "eosType " << eosType
         << " (AZH, CMF) is not part of this build; use eosType 0 (table), 1 (chiral) or 5 "
            "(ideal gas p=e/3)\n";
    return 0;
  }
  // hadronic EoS for the freeze-out surface (src/main.cpp:304-314)
  EoS *eosH = nullptr;
  if (eosTypeHadron == 0) {
    eosH = new EoSHadron("eos/eosHadronLog.dat");
  } else if (eosTypeHadron == -1) {
    eosH = eos;
  } else {
    cout << "eosTypeHadron " << eosTypeHadron
         << " (SMASH) is not part of this build; use 0 (PDG hadronic EoS, "
            "eos/eosHadronLog.dat) or -1 (the hydro EoS)\n";
    return 0;
  }
  if (train) {
    const vh_eos_t d = eos->device();
    return trainC2P(train, seed, &d);
  }
  TransportCoeff *trcoeff =
      new TransportCoeff(etaS, zetaS, ah, al, aRho, T0, etaSMin, etaSEpsilonMin, etaSShiftMuB,
                         etaSScaleMuB, zetaSPeakEpsilon, zetaSScaleBeta, zetaSSigmaMinus,
                         zetaSSigmaPlus, eos, etaSparam, zetaSparam);
  trcoeff->zetaSPeakT = zetaSPeakT;
  trcoeff->zetaSWidthMinus = zetaSWidthMinus;
  trcoeff->zetaSWidthPlus = zetaSWidthPlus;
  Fluid *f = new Fluid(eos, eosH, trcoeff, nx, ny, nz, xmin, xmax, ymin, ymax, etamin, etamax,
                       dtau, eCrit);
  if (precision != 0 || c2pMode != 0) {
    std::vector<float> model;
    if (c2pMode == VH_C2P_LEARNED) {
      model.resize(VH_MLP_NPARAM);
      if (c2pModel.empty() || !vhlle_learn::loadC2PModel(c2pModel.c_str(), model.data())) {
        cout << "c2pMode 1 needs c2pModel (a file written by vhlle -trainC2P)\n";
        return 1;
      }
    }
    f->configure(precision == 1 ? VH_PREC_MIXED : VH_PREC_FP64, c2pMode,
                 model.empty() ? nullptr : model.data());
  }
  cout << "fluid allocation done\n";
  if (icModel == 4) {
    ICGubser ic(gubserScale);
    ic.setIC(f, eos, tau0);
  } else if (icModel == 10) {
    setGaussianIC(f, eos, tau0);
  } else {
    cout << "icModel = " << icModel << " not implemented in this build (4, 10)\n";
    return 0;
  }
  cout << "IC done\n";
  Hydro *h = new Hydro(f, eos, trcoeff, tau0, dtau);
  if (vorticityOn) h->enableVorticity();  // src/main.cpp:380-383
  bool resized = false;
  f->initOutput(outputDir.c_str(), tau0, freezeoutOnly);  // src/main.cpp:389
  f->outputCorona(tau0, freezeoutExtend);  // src/main.cpp:389-391
  if (vorticityOn) f->printDbetaHeader();
  VtkOutput vtk_out(outputDir, eos, xmin, ymin, etamin);  // src/main.cpp:395-396
  long steps = 0;
  int nelements = 0;
  const auto t0 = chrono::steady_clock::now();
  do {
    if (!vtk_values.empty()) vtk_out.write(*h, vtk_values);  // src/main.cpp:402-404
    // sub-steps while dtau exceeds tau * d(eta) (src/main.cpp:405-419)
    int nSub = 1;
    while (dtau / nSub > 1.0 * h->getTau() * (etamax - etamin) / (nz - 1)) nSub *= 2;
    if (nSub > 1) {
      h->setDtau(h->getDtau() / nSub);
      h->performSteps(nSub);
      h->setDtau(h->getDtau() * nSub);
      cout << "timestep reduced by " << nSub << endl;
    } else {
      h->performStep();
    }
    steps += nSub;
    nelements = f->outputSurface(h->getTau(), freezeoutExtend);
    if (!freezeoutOnly) f->outputGnuplot(h->getTau());  // src/main.cpp:421-422
    if (h->getTau() >= tauResize && !resized) {  // src/main.cpp:423-427
      cout << "grid resize\n";
      // (the sub-step test keeps using the params' dtau, as the reference's)
      f = expandGrid2x(h, eos, eosH, trcoeff);
      resized = true;
    }
  } while (h->getTau() < tauMax + 0.0001 && (nelements > 0 || !stopOnEmptySurface));
  f->renameOutput(outputDir.c_str());  // src/main.cpp:435
  f->closeOutput();
  const double secs = chrono::duration<double>(chrono::steady_clock::now() - t0).count();
  cout << "tau = " << h->getTau() << "  steps = " << steps
       << "  sum Q^tau = " << totalEnergy(f, h->getTau()) << endl;
  cout << "Execution time = " << secs << " [sec], "
       << (double)f->ncell() * steps / secs << " cell-updates/s" << endl;
  if (dumpState) {
    f->syncFromDevice();
    const string fn = outputDir + "/state.bin";
    FILE *o = fopen(fn.c_str(), "wb");
    if (o) {
      double Q[7];
      std::vector<double> buf(7 * f->ncell());
      for (int iz = 0; iz < nz; iz++)
        for (int iy = 0; iy < ny; iy++)
          for (int ix = 0; ix < nx; ix++) {
            f->getCell(ix, iy, iz)->getQ(Q);
            const long c = ix + (long)nx * (iy + (long)ny * iz);
            for (int k = 0; k < 7; k++) buf[k * f->ncell() + c] = Q[k];
          }
      fwrite(buf.data(), sizeof(double), buf.size(), o);
      fclose(o);
      cout << "wrote " << fn << endl;
    }
  }
  delete h;
  delete f;
  delete trcoeff;
  if (eosH != eos) delete eosH;
  delete eos;
  return 0;
}
