/*
 * res.h -- result-buffer constructors (library-internal in the reference too).
 *
 * Mirrors nahmedraja/GASAL2 src/res.h:4-9 (src/res.cpp): which arrays exist
 * depends on algo / start_pos / secondBest exactly as in src/res.cpp:8-73.
 */
#ifndef __RES_H__
#define __RES_H__

#include "gasal.h"
#include "args_parser.h"

#ifdef __cplusplus
extern "C" {
#endif
/* src/res.cpp:8-73 */
gasal_res_t *gasal_res_new_host(uint32_t max_n_alns, Parameters *params);
/* src/res.cpp:76-101 */
gasal_res_t *gasal_res_new_device(gasal_res_t *device_cpy);
/* src/res.cpp:106-140 */
gasal_res_t *gasal_res_new_device_cpy(uint32_t max_n_alns, Parameters *params);
/* src/res.cpp:143-158 */
void gasal_res_destroy_host(gasal_res_t *res);
/* src/res.cpp:160-177 */
void gasal_res_destroy_device(gasal_res_t *device_res, gasal_res_t *device_cpy);
#ifdef __cplusplus
}
#endif

#endif
