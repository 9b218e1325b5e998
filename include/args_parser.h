/*
 * args_parser.h -- run-time options of an alignment batch.
 *
 * Drop-in for nahmedraja/GASAL2 src/args_parser.h:24-68: class Parameters with
 * the same public members, defaults (src/args_parser.cpp:12-31) and the test
 * program's command-line parser.
 */
#ifndef ARGS_PARSER_H
#define ARGS_PARSER_H

#include <string.h>
#include "gasal.h"

enum fail_type { NOT_ENOUGH_ARGS, TOO_MANY_ARGS, WRONG_ARG, WRONG_FILES, WRONG_ALGO };

#ifndef __cplusplus
/* C hosts hold a Parameters only through a pointer: they create it with
 * gasal2_parameters_new() and set its fields with gasal2_parameters_set()
 * (gasal2_ext.h). */
typedef enum fail_type fail_type;
typedef struct Parameters Parameters;
#else
#include <fstream>
#include <iostream>
#include <string>

class Parameters {
  public:
    Parameters(int argc, char **argv);
    ~Parameters();
    void print();
    void failure(fail_type f);
    void help();
    void parse();
    void fileopen();

    int32_t sa;
    int32_t sb;
    int32_t gapo;
    int32_t gape;
    comp_start start_pos;
    int print_out;
    int n_threads;
    int32_t k_band;

    Bool secondBest;

    bool isPacked;
    bool isReverseComplement;

    data_source semiglobal_skipping_head;
    data_source semiglobal_skipping_tail;

    algo_type algo;

    std::string query_batch_fasta_filename;
    std::string target_batch_fasta_filename;

    std::ifstream query_batch_fasta;
    std::ifstream target_batch_fasta;

  protected:
  private:
    int argc;
    char **argv;
};
#endif /* __cplusplus */

#endif
