// eqtl CLI over the MI355X engine (main.cpp:8-21 of the reference): errors
// print "[E] msg" and exit with code 1.
#include <exception>
#include <iostream>

#include "eqtl_host.hpp"

int main(int argc, char** argv) {
  try {
    eqtl::Options opt;
    const int rc = eqtl::parse_options(argc, argv, opt);
    if (rc == 2) {
      std::cout << "Program: eqtl (expression QTL mapping, MI355X engine)\n"
                   "Usage: eqtl -b <bfile> -e <pheno.tsv> [-c covar] [-g gff] [-k grm] [--model lm,lmm]\n"
                   "            [--mode trans|gw] [-w window] [--pval-trans p] [--precision fp32|bf16x3]\n"
                   "            [--delta ref|newton] [-o out]\n"
                   "       eqtl -b <bfile> --make-grm [-e pheno.tsv] -o <prefix>\n";
      return 0;
    }
    if (rc != 0) return 1;
    if (opt.make_grm) return eqtl::run_make_grm(opt);
    return eqtl::run_eqtl(opt);
  } catch (const std::exception&) {
    return 1;
  }
}
